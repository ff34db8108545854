#!/usr/bin/env python3
"""Throughput of the sparse hot path on one MI355X.

Default workload: power-law FP64 SpMV, y(i) = A(i,j) * x(j), A CSR
2,000,000 x 2,000,000 with Zipf(2.0) row lengths on [1, 1e5] rescaled to mean
16 (~32M entries), uniform distinct columns, values and x U[-1,1] (synthetic,
seed-generated).  A "step" is one compute() over that matrix with every input
already resident in HBM.

Prints one JSON line (metric, value, roofline, cpu_baseline, ...).  The CPU
baseline is taco's own generated OpenMP code (the real taco, compiled from the
reference sources into oracle/_ref) on the host cores, or the oracle's C
restatement of it when that library is absent; its output also checks the GPU
result.

Other workloads (--workload): spmm_f64, spmm_f32, spmm_bf16 (same A, dense B
2,000,000 x 64), spgemm (500,000^2, 8 entries per row), spmv_uniform (100,000^2,
10 entries per row), batched_spmv / batched_spmm (4,096 instances), selector,
codegen (generated modules beside hand-written kernels), gather_ceiling (random
8-byte load rate: the bound of the gather-limited SpMV kernels),
copy_calibration (scales the rocprofv3 traffic counters).
"""
import argparse
import json
import os
import sys
import time

# The CPU baseline's OpenMP threads are pinned one per core, packed; the
# OpenMP runtimes read these when they start, so they are set before any
# library that carries one (torch, the CPU oracle) is loaded.  Once an
# OpenMP runtime starts with binding on, it pins the calling thread to one
# core, so the process's CPU set is read here, before that.
try:
    AFFINITY_CPUS = len(os.sched_getaffinity(0))
except AttributeError:
    AFFINITY_CPUS = os.cpu_count() or 1
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0
SPMM_K = 64


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cgroup_cpu_quota():
    """CPUs granted by the cgroup's CPU quota (cgroup v2 cpu.max, v1
    cfs_quota/period), or None when unlimited or unreadable."""
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        if q != "max":
            return max(1, int(int(q) // int(per)))
        return None
    except (OSError, ValueError):
        pass
    try:
        with open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us") as f:
            q = int(f.read())
        with open("/sys/fs/cgroup/cpu/cpu.cfs_period_us") as f:
            per = int(f.read())
        return max(1, q // per) if q > 0 else None
    except (OSError, ValueError):
        return None


def affinity_cpus():
    return AFFINITY_CPUS


def host_cores():
    """Every host core this process may run on: the CPUs of its affinity
    mask, limited by the cgroup CPU quota (more threads than the quota
    would only be throttled)."""
    n = affinity_cpus()
    q = cgroup_cpu_quota()
    return max(1, min(n, q) if q else n)


def cpu_threading():
    return {"affinity_cpus": affinity_cpus(), "cgroup_quota_cpus": cgroup_cpu_quota(),
            "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
            "OMP_PLACES": os.environ.get("OMP_PLACES")}


def host_cpu():
    """CPU model and the cores this process may use (the baseline's host)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    out = {"model": model, "logical_cpus": os.cpu_count()}
    out.update(cpu_threading())
    return out


def trimmed_mean(ts):
    ts = sorted(ts)
    cut = int(len(ts) * 0.1)
    core = ts[cut:len(ts) - cut] if len(ts) - 2 * cut > 0 else ts
    return float(np.mean(core))


# ---------------------------------------------------------------- CPU leg --

def cpu_baseline_spmv(A, x, reps):
    """Times taco's CPU SpMV on the host (before the GPU is touched)."""
    import oracle
    cores = host_cores()
    if oracle.ref_available():
        y, ts = oracle.ref_spmv(A, x, threads=cores, reps=reps)
        kind = "reference"
    else:
        oracle.set_threads(cores)
        oracle.spmv(A, x)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            y = oracle.spmv(A, x)
            ts.append((time.perf_counter() - t0) * 1e3)
        kind = "port"
    return y, list(ts), cores, kind


def cpu_baseline_spmm(A, B, reps):
    import oracle
    cores = host_cores()
    if oracle.ref_available():
        C, ts = oracle.ref_spmm(A, B, threads=cores, reps=reps)
        kind = "reference"
    else:
        oracle.set_threads(cores)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            C = oracle.spmm(A, B)
            ts.append((time.perf_counter() - t0) * 1e3)
        kind = "port"
    return C, list(ts), cores, kind


def cpu_baseline_spgemm(A, B, reps):
    import oracle
    cores = host_cores()
    if oracle.ref_available():
        pos, crd, vals, ta, tc = oracle.ref_spgemm(A, B, threads=cores, reps=reps)
        kind = "reference"
        ts = [a + c for a, c in zip(ta, tc)]
    else:
        oracle.set_threads(cores)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            pos, crd = oracle.spgemm_assemble(A, B)
            vals = oracle.spgemm_compute(A, B, pos, crd)
            ts.append((time.perf_counter() - t0) * 1e3)
        kind = "port"
    return (pos, crd, vals), list(ts), cores, kind


# ---------------------------------------------------------------- GPU leg --

class Timer:
    """HIP events on torch's current stream, which the library launches on."""

    def __init__(self, torch):
        self.torch = torch
        self.e0 = torch.cuda.Event(enable_timing=True)
        self.e1 = torch.cuda.Event(enable_timing=True)

    def run(self, fn, steps, warmup, before_timed=None):
        torch = self.torch
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        if before_timed:
            before_timed()
        t0 = time.perf_counter()
        self.e0.record()
        for _ in range(steps):
            fn()
        self.e1.record()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3
        return self.e0.elapsed_time(self.e1) / steps, wall / steps


def gpu_setup():
    import torch
    from taco_amd import _lib
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    L = _lib.lib()
    _lib.check(L.taco_hip_set_stream(torch.cuda.current_stream().cuda_stream))
    return torch, L, _lib


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration recorded with it), or None when no summary
    covers that kernel on that workload."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        import json as _json
        with open(path) as f:
            data = _json.load(f)
        entry = data["kernels"][kernel]
        return entry["hbm_bytes_per_launch"] if entry["workload"] == workload else None
    except (OSError, KeyError, ValueError):
        return None


def sched(_lib, algo, block=0, items=0, group=0, rows=0):
    s = _lib.Schedule()
    s.algo, s.block_size, s.items_per_thread = algo, block, items
    s.group_size, s.rows_per_group = group, rows
    return s


def bench_spmv(args, A, x, name, workload_desc, spmm=None):
    """SpMV on one matrix.  spmm = spmm_inputs(...) adds the SpMM legs (f64,
    f32, bf16 with K = 64 on the same matrix) to the same JSON line."""
    import oracle
    from taco_amd import workloads as W
    y_ref, cpu_ts, cores, kind = (None, [], 0, None)
    if not args.no_cpu:
        log(f"cpu baseline: taco SpMV on {host_cores()} cores ...")
        y_ref, cpu_ts, cores, kind = cpu_baseline_spmv(A, x, args.cpu_reps)
    oracle.set_threads(host_cores())
    exact_y = oracle.exact_spmv(A, x)
    torch, L, _lib = gpu_setup()
    from taco_amd import descriptors as D
    pos = torch.from_numpy(A.pos).cuda()
    crd = torch.from_numpy(A.crd).cuda()
    vals = torch.from_numpy(A.vals).cuda()
    xd = torch.from_numpy(x).cuda()
    y = torch.empty(A.m, dtype=torch.float64, device="cuda")
    a, xdsc, ydsc = D.csr(pos, crd, vals, A.shape), D.dense(xd), D.dense(y)
    s_merge = sched(_lib, _lib.MERGE_PATH, 256, args.items)
    s_default = sched(_lib, _lib.THREAD_PER_ROW, 256)

    def step(s):
        return lambda: _lib.check(L.taco_hip_spmv_compute(ydsc.ptr, a.ptr, xdsc.ptr, s))

    timer = Timer(torch)
    # Slab-tiled plan: a one-time storage conversion of A, timed separately.
    # A plan of a small matrix is built first: the process's first plan
    # also pays for loading the build kernels' code (about 10 ms once per
    # process), which is no cost of this matrix's plan.
    small = W.csr_fixed(20000, 20000, 4, 1)
    sd = D.csr(torch.from_numpy(small.pos).cuda(), torch.from_numpy(small.crd).cuda(),
               torch.from_numpy(small.vals).cuda(), small.shape)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    warm = _lib.SpmvPlan(sd, 0, 0)
    torch.cuda.synchronize()
    first_plan_ms = (time.perf_counter() - t0) * 1e3
    del warm, sd
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = _lib.SpmvPlan(a, args.parts, 0)
    torch.cuda.synchronize()
    plan_ms = (time.perf_counter() - t0) * 1e3
    plan_info = plan.info()
    plan_shape = plan.shape()
    log(f"gpu: plan {plan_ms:.1f} ms {plan_info}; {args.warmup} warmup + {args.steps} steps ...")
    # Events around the main kernel of every timed step (same stream).
    ms, wall_ms = timer.run(lambda: plan.compute(ydsc, xdsc, args.unroll), args.steps, args.warmup,
                            before_timed=lambda: plan.time_kernel(args.steps))
    kernel_ms, kernel_launches = plan.kernel_time()
    plan.time_kernel(0)
    y_host = y.cpu().numpy()
    errors = {"slab_plan": fp64_error(y_host, exact_y, 1e-12)}
    ms_merge, _ = timer.run(step(s_merge), max(10, args.steps // 2), 3)
    y_merge = y.cpu().numpy()
    errors["merge_path"] = fp64_error(y_merge, exact_y, 1e-12)
    ms_default, _ = timer.run(step(s_default), max(5, args.steps // 10), 2)
    errors["default_thread_per_row"] = fp64_error(y.cpu().numpy(), exact_y, 1e-12)
    learned = None
    try:
        from taco_amd import selector as S
        from taco_amd import autotune
        sel = S.Selector()
        c, s_learned, pred_ms = sel.choose(a, 0)
        build_share = 0.0
        if s_learned is None:
            # the slab plan (the one built above); its build time is charged
            # to the PLAN_CALLS calls that reuse it, as the selector's
            # class times are
            run_learned = lambda: plan.compute(ydsc, xdsc)  # noqa: E731
            build_share = plan_ms / autotune.PLAN_CALLS
        else:
            run_learned = step(s_learned)
        ms_learned, _ = timer.run(run_learned, max(10, args.steps // 2), 3)
        errors["learned_" + S.CLASS_NAMES[c]] = fp64_error(y.cpu().numpy(), exact_y, 1e-12)
        learned = {"class": S.CLASS_NAMES[c], "ms": ms_learned,
                   "ms_with_plan_build_share": ms_learned + build_share,
                   "predicted_ms": pred_ms,
                   "speedup_vs_default": ms_default / (ms_learned + build_share)}
    except (OSError, ValueError) as e:  # no trained weights in this tree
        learned = {"error": str(e)}
    cold = None
    if A.nnz < 10_000_000:
        # Cold: a 512 MB write between launches evicts A and x from the MALL.
        flush = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        ts = []
        for _ in range(max(10, args.steps // 5)):
            flush.fill_(1.0)
            e0.record()
            plan.compute(ydsc, xdsc, args.unroll)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        cold = {"ms": float(np.median(ts)), "flush_bytes": flush.numel() * 8}
        del flush
    nbytes = W.algorithmic_bytes_spmv(A.m, A.n, A.nnz)
    gbs = nbytes / (ms * 1e-3) / 1e9
    check = None
    if y_ref is not None:
        # taco's CPU result against the FP64 oracle, and the GPU's against it
        scale = max(np.abs(y_ref).max(), 1e-300)
        errors["taco_cpu"] = fp64_error(y_ref, exact_y, 1e-12)
        check = {"slab_plan_vs_taco_cpu_normwise": float(np.abs(y_host - y_ref).max() / scale)}
    del exact_y
    cpu = None
    if cpu_ts:
        cms = trimmed_mean(cpu_ts)
        cpu = {"value": nbytes / (cms * 1e-3) / 1e9, "unit": "GB/s", "cores": cores, "host": host_cpu(),
               "kind": kind, "ms": cms, "median_ms": float(np.median(cpu_ts)),
               "sample": f"whole workload, {len(cpu_ts)} timed compute() calls "
                         f"(10% trimmed mean), JIT excluded, inputs in host DRAM"}
    traffic = profiled_traffic("k_slab_spmv_il", name)
    kernel_gbs = nbytes / (kernel_ms * 1e-3) / 1e9
    spmm_legs = None
    if spmm is not None:
        B, exact_c, exact_bf16, cpu_spmm = spmm
        del plan
        torch.cuda.empty_cache()
        spmm_legs = {}
        for dt in args.spmm_dtypes.split(","):
            log(f"gpu: SpMM {dt} K={B.shape[1]} ...")
            spmm_legs[dt] = spmm_leg(args, A, dt, pos, crd, B, exact_c, exact_bf16, timer,
                                     cpu_spmm if dt == "f64" else None)
    return {
        "metric": "SpMV HBM GB/s (% of 8 TB/s), SpMM GFLOP/s: learned vs default schedule at tol",
        "value": gbs, "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "wall_ms_per_step": wall_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seed-generated)",
        "config": {"workload": name, "desc": workload_desc, "m": A.m, "n": A.n,
                   "nnz": A.nnz, "schedule": f"slab-tiled plan parts={plan_info['parts']} "
                                             f"slab_width={plan_info['slab_width']} "
                                             f"unroll={args.unroll or plan_shape['unroll']} "
                                             f"threads={plan_shape['threads']}, interleaved "
                                             f"entries, segments ordered by column",
                   "parallelism": "single GPU"},
        "pct_of_hbm_peak": 100.0 * gbs / HBM_PEAK_GBS,
        "gflops": 2.0 * A.nnz / (ms * 1e-3) / 1e9,
        "nnz_per_s": A.nnz / (ms * 1e-3),
        "plan_build_ms": plan_ms,
        "first_plan_in_process_ms": first_plan_ms,
        "merge_path": {"ms": ms_merge, "gbs": nbytes / (ms_merge * 1e-3) / 1e9,
                       "note": "deterministic CSR merge-path kernel, no plan"},
        "default_schedule": {"schedule": "thread_per_row (taco split(i,256))",
                             "ms": ms_default, "speedup_vs_default": ms_default / ms},
        "learned_schedule": learned,
        "cold": cold,
        "roofline": {"bound": "hbm", "achieved": kernel_gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": kernel_gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_slab_spmv_il", "kernel_us": kernel_ms * 1e3,
                     "kernel_launches_timed": kernel_launches,
                     "algorithmic_bytes_per_launch": nbytes},
        "error_vs_fp64_oracle": errors,
        "check_vs_cpu_reference": check,
        "cpu_baseline": cpu,
        "spmm_k64": spmm_legs,
    }


def fp64_error(got, exact_pair, tol):
    """Condition-scaled error of a result against the FP64 oracle (gate
    |got - exact| <= tol * sum|a*b|, plus max elementwise-relative and
    normwise errors); computed outside every timed region."""
    import oracle
    exact, cond = exact_pair
    return oracle.error_report(got, exact, cond, tol)


SPMM_TOL = {"f64": 1e-12, "f32": 1e-5, "bf16": 1e-2}


def spmm_leg(args, A, dtype, pos, crd, B, exact, exact_bf16, timer, cpu=None):
    """SpMM C = A B with K = 64 in one precision form: the library's main
    schedule (merge path; MFMA for bf16), the learned selector's pick and
    taco's default schedule (one thread per row), each timed with HIP
    events; GFLOP/s, rooflines (algorithmic and gather-inclusive bytes) and
    the error of every result against the FP64 oracle."""
    import torch
    from taco_amd import _lib, descriptors as D, workloads as W
    L = _lib.lib()
    K = B.shape[1]
    if dtype == "bf16":
        vals = torch.from_numpy(A.vals.astype(np.float32)).to(torch.bfloat16).view(torch.int16).cuda()
        Bd = torch.from_numpy(B.astype(np.float32)).to(torch.bfloat16).view(torch.int16).cuda()
        C = torch.empty((A.m, K), dtype=torch.float32, device="cuda")
        cdt, sizes = "f32", (2, 2, 4)
    else:
        tdt = torch.float64 if dtype == "f64" else torch.float32
        vals = torch.from_numpy(A.vals).to(tdt).cuda()
        Bd = torch.from_numpy(B).to(tdt).cuda()
        C = torch.empty((A.m, K), dtype=tdt, device="cuda")
        cdt = dtype
        sizes = (8, 8, 8) if dtype == "f64" else (4, 4, 4)
    a = D.csr(pos, crd, vals, A.shape, dtype)
    bdsc, cdsc = D.dense(Bd, dtype), D.dense(C, cdt)
    main_name = "mfma" if dtype == "bf16" and args.mfma else f"merge{args.spmm_items}"
    s_main = sched(_lib, _lib.MFMA) if main_name == "mfma" else \
        sched(_lib, _lib.MERGE_PATH, 256, args.spmm_items, 64)
    s_default = sched(_lib, _lib.THREAD_PER_ROW, 256)

    def step(s_):
        return lambda: _lib.check(L.taco_hip_spmm_compute(cdsc.ptr, a.ptr, bdsc.ptr, s_))

    def result():
        torch.cuda.synchronize()
        return C.cpu().numpy().astype(np.float64)

    tol = SPMM_TOL[dtype]
    ms, wall_ms = timer.run(step(s_main), args.steps, args.warmup)
    errors = {main_name: fp64_error(result(), exact, tol)}
    if dtype == "bf16":
        errors[main_name + "_vs_bf16_rounded_inputs"] = fp64_error(result(), exact_bf16, 1e-5)
    learned = None
    try:
        from taco_amd import selector as S
        c, s_l, pred_ms = S.Selector().choose(a, 1, K)
        ms_l, _ = timer.run(step(s_l), max(5, args.steps // 2), 2)
        errors["learned_" + S.class_name(c)] = fp64_error(result(), exact, tol)
        learned = {"class": S.class_name(c), "ms": ms_l,
                   "predicted_ms": None if pred_ms != pred_ms else pred_ms,
                   "gflops": 2.0 * A.nnz * K / (ms_l * 1e-3) / 1e9}
    except (OSError, ValueError) as e:  # no trained weights in this tree
        learned = {"error": str(e)}
    ms_default, _ = timer.run(step(s_default), 2, 1)
    errors["default_thread_per_row"] = fp64_error(result(), exact, tol)
    if learned and "ms" in learned:
        learned["speedup_vs_default"] = ms_default / learned["ms"]
    nbytes = W.algorithmic_bytes_spmm(A.m, A.n, A.nnz, K, *sizes)
    gather_bytes = nbytes - sizes[1] * A.n * K + sizes[1] * A.nnz * K
    flops = 2.0 * A.nnz * K
    gbs = nbytes / (ms * 1e-3) / 1e9
    out = {
        "dtype": dtype if dtype != "bf16" else "bf16 (f32 accumulate)",
        "schedule": main_name, "ms": ms, "wall_ms": wall_ms,
        "gflops": flops / (ms * 1e-3) / 1e9,
        "learned_schedule": learned,
        "default_schedule": {"schedule": "thread_per_row (taco split(i,256))", "ms": ms_default,
                             "speedup_vs_default": ms_default / ms},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes": nbytes,
                     "gather_inclusive_bytes": gather_bytes,
                     "gather_inclusive_gbs": gather_bytes / (ms * 1e-3) / 1e9},
        "error_vs_fp64_oracle": errors,
    }
    if dtype == "bf16":
        # MFMA utilisation: flops the MFMA tiles issue vs the dense bf16 peak.
        out["mfma"] = {"useful_tflops": flops / (ms * 1e-3) / 1e12,
                       "peak_tflops": MFMA_BF16_PEAK_TFLOPS,
                       "frac": flops / (ms * 1e-3) / 1e12 / MFMA_BF16_PEAK_TFLOPS}
    if cpu is not None:
        out["cpu_baseline"] = cpu
    del a, bdsc, cdsc, vals, Bd, C
    torch.cuda.empty_cache()
    return out


def spmm_inputs(args, A, dtypes):
    """B (2,000,000 x 64, U[-1,1]), the FP64 oracle on it (and on its bf16
    rounding when bf16 is measured) and, for f64, taco's CPU SpMM timed on
    the host cores (a bounded sample of calls)."""
    import oracle
    from taco_amd import workloads as W
    K = SPMM_K
    B = W.uniform(A.n * K, 2, W.STREAM_B).reshape(A.n, K)
    cpu = None
    if not args.no_cpu and "f64" in dtypes:
        reps = max(2, args.cpu_reps // 6)
        log(f"cpu baseline: taco SpMM on {host_cores()} cores, {reps} calls ...")
        C_ref, cpu_ts, cores, kind = cpu_baseline_spmm(A, B, reps)
        cms = trimmed_mean(cpu_ts)
        cpu = {"value": 2.0 * A.nnz * K / (cms * 1e-3) / 1e9, "unit": "GFLOP/s", "cores": cores,
               "kind": kind, "ms": cms, "median_ms": float(np.median(cpu_ts)), "host": host_cpu(),
               "sample": f"whole workload, {len(cpu_ts)} timed compute() calls (10% trimmed mean)"}
        del C_ref
    oracle.set_threads(host_cores())
    log("fp64 oracle for SpMM ...")
    exact = oracle.exact_spmm(A, B)
    exact_bf16 = None
    if "bf16" in dtypes:
        import torch
        r = lambda v: torch.from_numpy(np.ascontiguousarray(v, dtype=np.float32)) \
            .to(torch.bfloat16).to(torch.float64).numpy()
        from taco_amd import workloads as W2
        Ar = W2.Csr(A.m, A.n, A.pos, A.crd, r(A.vals))
        exact_bf16 = oracle.exact_spmm(Ar, r(B))
    return B, exact, exact_bf16, cpu


def bench_spmm(args, A, dtype):
    """One SpMM precision form on its own JSON line (--workload spmm_<dtype>)."""
    B, exact, exact_bf16, cpu = spmm_inputs(args, A, (dtype,))
    torch, L, _lib = gpu_setup()
    pos = torch.from_numpy(A.pos).cuda()
    crd = torch.from_numpy(A.crd).cuda()
    leg = spmm_leg(args, A, dtype, pos, crd, B, exact, exact_bf16, Timer(torch),
                   cpu if dtype == "f64" else None)
    out = {
        "metric": "SpMM GFLOP/s", "value": leg["gflops"], "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": leg["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": leg["dtype"], "data": "synthetic (seed-generated)",
        "config": {"workload": f"spmm_powerlaw_2m_k64_{dtype}", "m": A.m, "nnz": A.nnz,
                   "k": SPMM_K, "schedule": leg["schedule"]},
        "roofline": leg["roofline"], "cpu_baseline": leg.get("cpu_baseline"),
    }
    out.update({k: v for k, v in leg.items() if k not in out})
    return out


def bench_spgemm(args):
    from taco_amd import workloads as W
    log("generating 500k x 500k, 8 per row ...")
    A, B = W.spgemm_500k()
    ref, cpu_ts, cores, kind = (None, [], 0, None)
    if not args.no_cpu:
        log("cpu baseline: taco two-phase SpGEMM ...")
        ref, cpu_ts, cores, kind = cpu_baseline_spgemm(A, B, max(2, args.cpu_reps // 3))
    torch, L, _lib = gpu_setup()
    import ctypes
    from taco_amd import descriptors as D
    ad = D.csr(torch.from_numpy(A.pos).cuda(), torch.from_numpy(A.crd).cuda(),
               torch.from_numpy(A.vals).cuda(), A.shape)
    bd = D.csr(torch.from_numpy(B.pos).cuda(), torch.from_numpy(B.crd).cuda(),
               torch.from_numpy(B.vals).cuda(), B.shape)
    outs = []

    def assemble():
        c = D.csr_empty((A.m, B.n))
        _lib.check(L.taco_hip_spgemm_assemble(c.ptr, ad.ptr, bd.ptr))
        outs.append(c)

    def release():
        for c in outs[:-1]:
            for p in (c.level_ptr(1, 0), c.level_ptr(1, 1), c.vals_ptr()):
                L.taco_hip_unified_free(ctypes.c_void_p(p))
        del outs[:-1]

    timer = Timer(torch)
    ms_asm = []
    for _ in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        assemble()
        torch.cuda.synchronize()
        ms_asm.append((time.perf_counter() - t0) * 1e3)
        release()
    ms_asm = float(np.median(ms_asm[args.warmup:]))
    c = outs[-1]
    ms_cmp, _ = timer.run(lambda: _lib.check(L.taco_hip_spgemm_compute(c.ptr, ad.ptr, bd.ptr)),
                          args.steps, args.warmup)
    nnzc = c.t.vals_size
    check = None
    if ref is not None:
        pos = np.empty(A.m + 1, np.int32)
        crd = np.empty(nnzc, np.int32)
        vals = np.empty(nnzc)
        for h, p in ((pos, c.level_ptr(1, 0)), (crd, c.level_ptr(1, 1)), (vals, c.vals_ptr())):
            _lib.check(L.taco_hip_memcpy(h.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p), h.nbytes, 1))
        check = {"pos_equal": bool(np.array_equal(pos, ref[0])),
                 "crd_equal": bool(np.array_equal(crd, ref[1])),
                 "vals_max_rel": float(np.abs(vals - ref[2]).max() / np.abs(ref[2]).max())}
    asm_bytes = 8 * (A.m + 1) + 4 * A.nnz + 4 * (B.m + 1) + 4 * B.nnz + 4 * nnzc
    cmp_bytes = 12 * A.nnz + 12 * B.nnz + 12 * (A.m + 1) + 12 * nnzc
    cpu = None
    if cpu_ts:
        cms = trimmed_mean(cpu_ts)
        cpu = {"value": nnzc / (cms * 1e-3), "unit": "nnz(C)/s", "cores": cores, "host": host_cpu(),
               "kind": kind, "ms": cms, "sample": "whole workload, assemble + compute"}
    return {
        "metric": "SpGEMM nnz(C)/s (assemble + compute)", "unit": "nnz(C)/s",
        "value": nnzc / ((ms_asm + ms_cmp) * 1e-3), "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_asm + ms_cmp, "assemble_ms": ms_asm,
        "compute_ms": ms_cmp, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seed-generated)",
        "config": {"workload": "spgemm_500k_8perrow", "nnz_c": nnzc},
        "roofline": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                     "achieved": cmp_bytes / (ms_cmp * 1e-3) / 1e9,
                     "frac": cmp_bytes / (ms_cmp * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": profiled_traffic("k_spgemm_numeric_short", "spgemm_500k_8perrow"),
                     "kernel": "k_spgemm_numeric_short",
                     "assemble_gbs": asm_bytes / (ms_asm * 1e-3) / 1e9},
        "check_vs_cpu_reference": check,
        "cpu_baseline": cpu,
    }


def bench_batched(args, op):
    """4,096 uneven CSR instances (rows log-uniform in [1e3, 2e5], row
    lengths uniform in [2, 64], mixed uniform / banded / 8x8-block
    structure, f32) in one persistent hipGraph-replayed launch: taco's
    default (one lane per row for SpMV, a 32-lane group per row for SpMM)
    against the exhaustive per-instance autotune."""
    torch, L, _lib = gpu_setup()
    from taco_amd import autotune, batch as BT
    t0 = time.perf_counter()
    wl = BT.BatchedWorkload(ninst=args.instances, rmin=1000, rmax=200000, kmin=2, kmax=64,
                            seed=4, width=32, spmm=(op == 1))
    gen_s = time.perf_counter() - t0
    log(f"batched: {wl.ninst} instances, {wl.total_rows} rows, {wl.total_nnz} entries, "
        f"generated in {gen_s:.1f}s")
    t0 = time.perf_counter()
    if args.batched_sched is None:
        best, times, cands, rejected = autotune.tune_batch(wl, op, reps=3)
        race = autotune.tune_batch.last_race
        runs = (("default", BT.default_schedule(op, wl.ninst)), ("autotuned", best))
    else:
        # one schedule for every instance, no autotuning (for profiling the
        # whole-batch kernel alone)
        best = np.full(wl.ninst, args.batched_sched, np.int32)
        rejected, race = [], None
        runs = (("autotuned", best),)
    tune_s = time.perf_counter() - t0
    timer = Timer(torch)
    res = {}
    for name, sched in runs:
        b = BT.Batch(wl, op, sched)
        b.capture()
        ms, wall = timer.run(b.launch_graph, args.steps, args.warmup)
        res[name] = {"ms": ms, "wall_ms": wall, "info": b.info()}
        del b
    nbytes = wl.algorithmic_bytes(op)
    flops = 2.0 * wl.total_nnz * (1 if op == 0 else wl.width)
    ms = res["autotuned"]["ms"]
    gbs = nbytes / (ms * 1e-3) / 1e9
    lanes = BT.SPMV_LANES if op == 0 else BT.SPMM_LANES
    return {
        "metric": f"batched {'SpMV' if op == 0 else 'SpMM'} GB/s (one persistent graph launch)",
        "value": gbs, "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seed-generated)",
        "config": {"workload": f"batched_{'spmv' if op == 0 else 'spmm_k32'}_4096",
                   "instances": wl.ninst, "rows": wl.total_rows, "nnz": wl.total_nnz,
                   "schedule": ("lanes per row from the exhaustive autotuner (per instance, "
                                "raced against uniform schedules on the whole batch)"
                                if args.batched_sched is None else
                                f"{args.batched_sched} lanes per row for every instance "
                                "(0 = stream), no autotuning")},
        "gflops": flops / (ms * 1e-3) / 1e9, "nnz_per_s": wl.total_nnz / (ms * 1e-3),
        "default_schedule": ({"ms": res["default"]["ms"],
                              "speedup_vs_default": res["default"]["ms"] / ms}
                             if "default" in res else None),
        "autotune": {"seconds": tune_s, "rejected": rejected,
                     "chosen_lanes_histogram": {int(l): int((best == l).sum()) for l in lanes},
                     "whole_batch_race_ms": race},
        "generate_seconds": gen_s,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_batch_spmv" if op == 0 else "k_batch_spmm",
                     "algorithmic_bytes_per_launch": nbytes},
        "cpu_baseline": None,
    }


def bench_selector(args):
    """The learned schedule selector on held-out problems (seed 8, never
    trained on): each problem is regenerated, every class re-timed on this
    GPU (the exhaustive autotune), and the selector's pick compared with the
    exhaustive best and with taco's default schedule (thread per row).
    Also times one decision (features + network) on the 2M-row power-law
    matrix against the SpMV it chooses."""
    torch, L, _lib = gpu_setup()
    from taco_amd import selector as S, descriptors as D, workloads as W
    sel = S.Selector()
    F, T, ops = [], [], []
    t0 = time.perf_counter()
    for i in range(args.problems):
        A, op, K, d = S.make_problem(8, i)
        prob = S.DeviceProblem(A, op, K, seed=i)
        F.append(S.features(prob.a, op, K))
        T.append(S.time_classes(prob, reps=5))
        ops.append(op)
        del prob, A
        if (i + 1) % 25 == 0:
            log(f"selector: {i + 1}/{args.problems} problems timed")
    F, T, ops = np.array(F), np.array(T), np.array(ops)
    tune_s = time.perf_counter() - t0
    choice, _ = sel.choose_from_features(F)
    ev = S.evaluate(choice, T, ops)
    per_op = {}
    for o, name in ((0, "spmv"), (1, "spmm")):
        k = ops == o
        if k.any():
            per_op[name] = S.evaluate(choice[k], T[k], ops[k])
    # decision latency on the power-law matrix
    A = W.powerlaw_2m()
    a = D.csr(torch.from_numpy(A.pos).cuda(), torch.from_numpy(A.crd).cuda(),
              torch.from_numpy(A.vals).cuda(), A.shape)
    lat = []
    for _ in range(args.warmup + 20):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        c, sched_c, pred_ms = sel.choose(a, 0)
        lat.append((time.perf_counter() - t1) * 1e3)
    lat = float(np.median(lat[args.warmup:]))
    x = D.dense(torch.from_numpy(W.uniform(A.n, 2, W.STREAM_X)).cuda())
    y = D.dense(torch.empty(A.m, dtype=torch.float64, device="cuda"))
    plan_ms = None
    if sched_c is None:
        # the slab plan: built once, then its compute is what the class runs
        from taco_amd import autotune
        plan, plan_ms = autotune.slab_plan(a)
        kms, _ = Timer(torch).run(lambda: plan.compute(y, x), 20, 3)
    else:
        kms, _ = Timer(torch).run(
            lambda: _lib.check(L.taco_hip_spmv_compute(y.ptr, a.ptr, x.ptr, sched_c)), 20, 3)
    return {
        "metric": "learned-schedule speedup over the default schedule (held-out problems)",
        "value": ev["default_over_learned"], "unit": "x", "n_gpus": 1, "steps": args.problems,
        "warmup": args.warmup, "ms_per_step": ev["learned_ms"] / len(F),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 network; f64 SpMV, f32 SpMM", "data": "synthetic (seed-generated)",
        "config": {"workload": "selector_heldout", "problems": int(len(F)),
                   "rows": "log-uniform 10k-2M", "structures": list(S.STRUCTURES),
                   "spmm_widths": list(S.SPMM_WIDTHS)},
        "learned_over_exhaustive": ev["learned_over_exhaustive"],
        "evaluation": ev, "per_op": per_op, "retime_seconds": tune_s,
        "decision": {"latency_ms": lat, "class": S.CLASS_NAMES[c], "predicted_ms": pred_ms,
                     "measured_ms": kms, "plan_build_ms": plan_ms, "matrix": "powerlaw_2m"},
        "roofline": None, "cpu_baseline": None,
    }


def bench_codegen(args):
    """Modules emitted by CodeGen_HIP (the reference's GPU schedules) beside
    the hand-written kernels of the same schedule family, on the 2M-row
    power-law matrix: SpMV f64 and SpMM f64 with K = 64."""
    torch, L, _lib = gpu_setup()
    from taco_amd import codegen, descriptors as D, workloads as W
    A = W.powerlaw_2m()
    x = W.uniform(A.n, 2, W.STREAM_X)
    pos = torch.from_numpy(A.pos).cuda()
    crd = torch.from_numpy(A.crd).cuda()
    vals = torch.from_numpy(A.vals).cuda()
    a = D.csr(pos, crd, vals, A.shape)
    xd = D.dense(torch.from_numpy(x).cuda())
    y = D.dense(torch.empty(A.m, dtype=torch.float64, device="cuda"))
    K = 64
    Bt = torch.rand((A.n, K), dtype=torch.float64, device="cuda")
    Ct = torch.empty((A.m, K), dtype=torch.float64, device="cuda")
    b, c = D.dense(Bt), D.dense(Ct)
    timer = Timer(torch)
    rows = {}
    for name in ("spmv_default_f64", "spmv_rows_f64", "spmv_rows64_f64", "spmv_nnz_f64",
                 "spmv_merge_f64", "spmv_block_f64", "spmm_default_f64", "spmm_nnz_f64",
                 "spmm_merge_f64"):
        k = codegen.Kernel(codegen.standard(name))
        descs = (y, a, xd) if name.startswith("spmv") else (c, a, b)
        reps = 3 if "default" in name else args.steps
        ms, _ = timer.run(lambda: k.compute(*descs), reps, 1)
        rows[name] = {"ms": ms}
    hand = {"spmv_thread_per_row": sched(_lib, _lib.THREAD_PER_ROW, 256),
            "spmv_group32": sched(_lib, _lib.GROUP_PER_ROW, 256, group=32),
            "spmv_nnz_split8": sched(_lib, _lib.NNZ_SPLIT, 256, 8),
            "spmv_merge8": sched(_lib, _lib.MERGE_PATH, 256, 8)}
    for name, s_ in hand.items():
        ms, _ = timer.run(lambda: _lib.check(L.taco_hip_spmv_compute(y.ptr, a.ptr, xd.ptr, s_)),
                          3 if "thread" in name else args.steps, 1)
        rows["hand_" + name] = {"ms": ms}
    s_mm = sched(_lib, _lib.MERGE_PATH, 256, 256, 64)
    ms, _ = timer.run(lambda: _lib.check(L.taco_hip_spmm_compute(c.ptr, a.ptr, b.ptr, s_mm)),
                      args.steps, 1)
    rows["hand_spmm_merge256"] = {"ms": ms}
    nb = W.algorithmic_bytes_spmv(A.m, A.n, A.nnz)
    for k_, v in rows.items():
        if "spmv" in k_:
            v["gbs"] = nb / (v["ms"] * 1e-3) / 1e9
        else:
            v["gflops"] = 2.0 * A.nnz * K / (v["ms"] * 1e-3) / 1e9
    best = rows["spmv_merge_f64"]
    return {
        "metric": "CodeGen_HIP modules vs hand-written kernels (power-law 2M)",
        "value": best["gbs"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": best["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seed-generated)",
        "config": {"workload": "codegen_powerlaw_2m", "nnz": A.nnz, "k": K},
        "kernels": rows, "roofline": None, "cpu_baseline": None,
    }


# nnz of the 2M-row power-law matrix: one x gather per entry
POWERLAW_GATHERS = 31_286_882


def gather_ceiling(args):
    """The rate at which MI355X serves random 8-byte loads with nothing else
    in flight (taco_hip_gather_ceiling), swept over the footprint of the
    gathered array, loads in flight per lane, workgroups per CU, and the
    vector-L1 / scalar-cache load paths.  Each point reports the time the
    power-law SpMV's 31.3M x gathers would take at that rate: the ceiling the
    slab kernel (one gather per entry, 1 MB x slabs) is measured against."""
    import ctypes
    torch, L, _lib = gpu_setup()
    out = (ctypes.c_double * 3)()
    reps = max(3, min(args.steps, 20))
    rows = []
    if args.ceiling_points:
        # log2:mode:unroll:wg_per_cu points only (for counter runs)
        points = [tuple(int(v) for v in p.split(":")) for p in args.ceiling_points.split(",")]
    else:
        points = [(log2, mode, u, wg) for log2 in (11, 14, 17, 21, 25)
                  for mode, unrolls, wgs in ((0, (1, 4, 8, 16), (4, 8)), (1, (4, 16), (8,)),
                                             (2, (4, 8), (8,)))
                  for u in unrolls for wg in wgs]
    for log2, mode, u, wg in points:
        log(f"gather ceiling: footprint 2^{log2} f64, mode {mode}, unroll {u}, "
            f"{wg} workgroups/CU")
        _lib.check(L.taco_hip_gather_ceiling(mode, u, log2, wg, POWERLAW_GATHERS, reps, out),
                   "gather ceiling")
        ms, lane_loads, wave_loads = out[0], out[1], out[2]
        rate = (lane_loads + wave_loads) / (ms * 1e-3)
        rows.append({"footprint_bytes": 8 << log2, "path": ("vector", "scalar", "mixed")[mode],
                     "unroll": u, "wg_per_cu": wg, "ms": ms, "lane_loads": lane_loads,
                     "wave_loads": wave_loads, "gloads_per_s": rate / 1e9,
                     "powerlaw_gathers_us": POWERLAW_GATHERS / rate * 1e6})
    best = {}
    for r in rows:
        k = (r["footprint_bytes"], r["path"])
        if k not in best or r["gloads_per_s"] > best[k]["gloads_per_s"]:
            best[k] = r
    l2 = best.get((8 << 17, "vector")) or max(rows, key=lambda r: r["gloads_per_s"])
    return {
        "metric": "random 8-byte gather rate (G loads/s), 1 MB footprint, vector path",
        "value": l2["gloads_per_s"], "unit": "G loads/s", "n_gpus": 1, "steps": reps,
        "warmup": 1, "ms_per_step": l2["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (in-register index stream)",
        "config": {"workload": "gather_ceiling", "gathers": POWERLAW_GATHERS},
        "best": [{"footprint_bytes": k[0], "path": k[1], **{kk: v[kk] for kk in
                  ("unroll", "wg_per_cu", "gloads_per_s", "powerlaw_gathers_us")}}
                 for k, v in sorted(best.items())],
        "sweep": rows, "roofline": None, "cpu_baseline": None,
    }


def copy_calibration(args):
    """A plain device copy of known size (2 GiB read + 2 GiB written per
    launch), run under rocprofv3 --pmc to scale FETCH_SIZE / WRITE_SIZE into
    HBM bytes for the same counters read on the sparse kernels."""
    torch, _, _ = gpu_setup()
    nbytes = 1 << 31
    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").uniform_()
    dst = torch.empty_like(src)
    timer = Timer(torch)
    ms, _ = timer.run(lambda: dst.copy_(src), args.steps, args.warmup)
    return {"metric": "device copy GB/s", "value": 2 * nbytes / (ms * 1e-3) / 1e9,
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "copy_calibration_2GiB", "read_bytes": nbytes,
                       "write_bytes": nbytes}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="spmv_powerlaw")
    ap.add_argument("--items", type=int, default=8, help="SpMV merge items per thread")
    ap.add_argument("--parts", type=int, default=0, help="slab plan parts (0 = library default)")
    ap.add_argument("--unroll", type=int, default=0, help="slab plan entries per thread per step (0 = library default)")
    ap.add_argument("--spmm-items", type=int, default=256)
    ap.add_argument("--mfma", type=int, default=1)
    ap.add_argument("--spmm-dtypes", default="f64,f32,bf16",
                    help="SpMM legs of the default workload (empty: SpMV only)")
    ap.add_argument("--cpu-reps", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--instances", type=int, default=4096, help="batched workloads")
    ap.add_argument("--batched-sched", type=int, default=None,
                    help="batched workloads: one lane count for every instance, no autotuning")
    ap.add_argument("--problems", type=int, default=200, help="selector held-out problems")
    ap.add_argument("--ceiling-points", default="",
                    help="gather_ceiling: only these log2:mode:unroll:wg_per_cu points")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU runs are out of scope for this backend)", file=sys.stderr)
        sys.exit(2)
    from taco_amd import workloads as W
    w = args.workload
    if w in ("spmv_powerlaw", "spmm_f64", "spmm_f32", "spmm_bf16"):
        log("generating 2M x 2M power-law matrix ...")
        t0 = time.perf_counter()
        A = W.powerlaw_2m()
        log(f"  nnz={A.nnz} in {time.perf_counter() - t0:.1f}s")
        if w == "spmv_powerlaw":
            x = W.uniform(A.n, 2, W.STREAM_X)
            spmm = None
            if args.spmm_dtypes:
                spmm = spmm_inputs(args, A, args.spmm_dtypes.split(","))
            out = bench_spmv(args, A, x, "spmv_powerlaw_2m_f64",
                             "2Mx2M CSR, Zipf(2) rows on [1,1e5] rescaled to mean 16",
                             spmm=spmm)
        else:
            out = bench_spmm(args, A, w.split("_")[1])
    elif w == "spmv_uniform":
        A, x = W.spmv_uniform_100k()
        out = bench_spmv(args, A, x, "spmv_uniform_100k_f64", "100kx100k, 10 per row")
    elif w == "spgemm":
        out = bench_spgemm(args)
    elif w == "copy_calibration":
        out = copy_calibration(args)
    elif w == "gather_ceiling":
        out = gather_ceiling(args)
    elif w == "selector":
        out = bench_selector(args)
    elif w == "codegen":
        out = bench_codegen(args)
    elif w in ("batched_spmv", "batched_spmm"):
        out = bench_batched(args, 0 if w == "batched_spmv" else 1)
    else:
        sys.exit(f"unknown workload {w}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
