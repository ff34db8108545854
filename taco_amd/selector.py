"""Learned schedule selector for single-matrix SpMV and SpMM.

A small MLP maps 32 sparsity features of a CSR matrix (computed on the GPU
by ``taco_hip_features``) to one of 24 schedule classes — the 13 SpMV
schedules of ``autotune.SPMV_CANDIDATES`` and the 11 SpMM schedules of
``autotune.spmm_candidates`` — plus a predicted log time.  The network is
trained on the GPU by the library's own forward, backward and Adam kernels
(``taco_hip_mlp_*``) on timings the exhaustive autotuner measures; the
autotuner stays its fallback and arbiter.  The selector never changes a
result: it only picks which of the implemented schedules runs, and every
schedule computes the same values to the value tolerance.

No reference counterpart: taco's schedules are written by hand in the
scheduling tests (test/tests-scheduling-eval.cpp:193-369).

Training data (``make_problem``): square matrices, rows log-uniform in
[10k, 2M], one of four structures — uniform random columns with uniform row
lengths, Zipf row lengths (alpha in [1.5, 3]) with uniform columns, banded,
or block diagonal — SpMV in f64 or SpMM in f32 with width K in
{16, 32, 64, 128, 256} (rows capped so that B holds at most 2^25 values).
"""
import ctypes
import os

import numpy as np

from . import _lib
from . import autotune
from . import workloads as W

N_FEATURES = 32
N_CLASSES = 24
N_SPMV = 13
N_OUT = 25
SPMM_WIDTHS = (16, 32, 64, 128, 256)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                               "selector.npz")

SPMV_NAMES = [name for name, _ in autotune.SPMV_CANDIDATES]
SPMM_NAMES = [name for name, _ in autotune.spmm_candidates(64)]
CLASS_NAMES = SPMV_NAMES + ["spmm_" + n for n in SPMM_NAMES]
assert len(SPMV_NAMES) == N_SPMV and len(CLASS_NAMES) == N_CLASSES
# The untuned schedule (taco's default GPU schedule: one thread per row).
DEFAULT_CLASS = {0: SPMV_NAMES.index("thread_per_row"),
                 1: N_SPMV + SPMM_NAMES.index("thread_per_row")}


# Classes whose kernels accumulate with atomics (the nnz-split schedules and
# the slab plan's LDS row partials): their results can differ in the last
# bits from run to run, always within the value tolerance.  Every other
# class sums in a fixed order.
ATOMIC_CLASSES = {c for c, name in enumerate(CLASS_NAMES)
                  if "nnz_split" in name or name == "slab_plan"}
# The slab plan (a storage conversion of A, reused across calls): its class
# time is the per-call time plus build time / autotune.PLAN_CALLS.
PLAN_CLASS = SPMV_NAMES.index("slab_plan")


# bf16 SpMM goes to the MFMA kernel by rule: the network's classes were
# timed in f64 and f32, the MFMA kernel exists only for bf16 (K a multiple of
# 4), and it is the fastest bf16 form measured (0.89 ms on the 2M-row
# power-law matrix at K = 64, against 2.06 ms for the merge class the
# network picks there).  Its index lies past the network's classes.
RULE_MFMA = N_CLASSES


def class_range(op):
    return (0, N_SPMV) if op == 0 else (N_SPMV, N_CLASSES)


def class_name(c):
    return "spmm_mfma_bf16 (rule)" if c == RULE_MFMA else CLASS_NAMES[c]


def class_schedule(c, K=0):
    """The taco_hip_schedule of class c (SpMM group sizes follow K); None
    for the slab plan."""
    if c == RULE_MFMA:
        return autotune._sched(_lib.MFMA)
    if c < N_SPMV:
        return autotune.SPMV_CANDIDATES[c][1]
    return autotune.spmm_candidates(K)[c - N_SPMV][1]


def mfma_rule(a_desc, op, K):
    """True when the bf16 rule applies: SpMM over bf16 values, K % 4 == 0."""
    return op == 1 and int(a_desc.t.csize) == 16 and K > 0 and K % 4 == 0


# ------------------------------------------------------------ features --

def csr_stats(a_desc):
    s = (ctypes.c_int64 * 30)()
    _lib.check(_lib.lib().taco_hip_csr_stats(a_desc.ptr, s), "csr stats")
    return np.array(s, dtype=np.int64)


def features_from_stats(stats, op, K):
    s = np.ascontiguousarray(stats, dtype=np.int64)
    f = np.zeros(N_FEATURES, dtype=np.float32)
    _lib.check(_lib.lib().taco_hip_features_from_stats(
        s.ctypes.data_as(ctypes.c_void_p), op, K, f.ctypes.data_as(ctypes.c_void_p)))
    return f


def features(a_desc, op, K=0):
    f = np.zeros(N_FEATURES, dtype=np.float32)
    _lib.check(_lib.lib().taco_hip_features(a_desc.ptr, op, K,
                                            f.ctypes.data_as(ctypes.c_void_p)), "features")
    return f


# ------------------------------------------------------------- network --

class Mlp:
    """Device MLP 32 -> 128 -> 128 -> 25 driven through the C ABI."""

    def __init__(self, params=None, seed=0):
        L = _lib.lib()
        self._p = ctypes.c_void_p()
        _lib.check(L.taco_hip_mlp_create(ctypes.byref(self._p)), "mlp create")
        self.n_params = int(L.taco_hip_mlp_param_count())
        self.set_params(init_params(seed) if params is None else params)

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.n_params
        _lib.check(_lib.lib().taco_hip_mlp_set_params(self._p, p.ctypes.data_as(ctypes.c_void_p)))

    def params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        _lib.check(_lib.lib().taco_hip_mlp_get_params(self._p, p.ctypes.data_as(ctypes.c_void_p)))
        return p

    def forward(self, X):
        """X: float32 cuda tensor [B, 32] -> float32 cuda tensor [B, 25]."""
        import torch
        Z = torch.empty((X.shape[0], N_OUT), dtype=torch.float32, device=X.device)
        _lib.check(_lib.lib().taco_hip_mlp_forward(self._p, ctypes.c_void_p(X.data_ptr()),
                                                   X.shape[0], ctypes.c_void_p(Z.data_ptr())))
        return Z

    def gradients(self, X, cls, logt):
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = ctypes.c_float()
        _lib.check(_lib.lib().taco_hip_mlp_gradients(
            self._p, ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(cls.data_ptr()),
            ctypes.c_void_p(logt.data_ptr()), X.shape[0], g.ctypes.data_as(ctypes.c_void_p),
            ctypes.byref(loss)), "mlp gradients")
        return float(loss.value), g

    def train_step(self, X, cls, logt, lr, weight_decay=0.0, want_loss=False):
        loss = ctypes.c_float()
        _lib.check(_lib.lib().taco_hip_mlp_train_step(
            self._p, ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(cls.data_ptr()),
            ctypes.c_void_p(logt.data_ptr()), X.shape[0], lr, weight_decay,
            ctypes.byref(loss) if want_loss else None), "mlp train step")
        return float(loss.value) if want_loss else None

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None and _lib._lib is not None:
            _lib._lib.taco_hip_mlp_destroy(self._p)
            self._p = None


def init_params(seed):
    """He-uniform weights, zero biases, in the flat layout of the C ABI."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((32, 128), (128, 128), (128, 25)):
        lim = np.sqrt(6.0 / fan_in)
        parts.append(rng.uniform(-lim, lim, size=(fan_out, fan_in)).ravel())
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


# ------------------------------------------------------------ problems --

STRUCTURES = ("uniform", "zipf", "banded", "block")


def problem_op(seed, index):
    """The operation (0 SpMV, 1 SpMM) make_problem(seed, index) draws."""
    return int(np.random.default_rng([seed, index]).integers(0, 2))


def make_problem(seed, index, rmin=10_000, rmax=2_000_000, op=None):
    """A seeded random problem: (host Csr, op, K, description dict)."""
    rng = np.random.default_rng([seed, index])
    drawn = int(rng.integers(0, 2))
    op = drawn if op is None else op
    K = int(rng.choice(SPMM_WIDTHS)) if op == 1 else 0
    hi = rmax if op == 0 else max(rmin, min(rmax, (1 << 25) // K))
    m = int(np.exp(rng.uniform(np.log(rmin), np.log(hi))))
    structure = STRUCTURES[int(rng.integers(0, 4))]
    mseed = int(rng.integers(1, 1 << 40))
    desc = {"m": m, "op": op, "K": K, "structure": structure}
    if structure == "uniform":
        kmax = int(np.exp(rng.uniform(np.log(2), np.log(96))))
        kmin = int(rng.integers(0, kmax + 1))
        A = W.csr_from_counts(m, m, W.uniform_counts(m, kmin, kmax, mseed), mseed)
        desc.update(kmin=kmin, kmax=kmax)
    elif structure == "zipf":
        alpha = float(rng.uniform(1.5, 3.0))
        mean = float(np.exp(rng.uniform(np.log(2), np.log(48))))
        A = W.csr_powerlaw(m, m, mseed, alpha=alpha, kmax=min(100_000, m), mean=mean)
        desc.update(alpha=alpha, mean=mean)
    elif structure == "banded":
        band = int(rng.choice([8, 16, 32, 64, 128, 256]))
        kmax = int(rng.integers(2, min(band + 1, 96) + 1))
        A = W.csr_from_counts(m, m, W.uniform_counts(m, 1, kmax, mseed), mseed,
                              W.STRUCT_BANDED, band)
        desc.update(band=band, kmax=kmax)
    else:
        b = int(rng.choice([4, 8, 16, 32, 64]))
        rows = np.arange(m, dtype=np.int64)
        full = np.minimum(b, m - (rows // b) * b).astype(np.int32)
        fill = float(rng.uniform(0.3, 1.0))
        counts = np.maximum(1, np.floor(full * fill)).astype(np.int32)
        A = W.csr_from_counts(m, m, counts, mseed, W.STRUCT_BLOCK, b)
        desc.update(block=b, fill=fill)
    return A, op, K, desc


class DeviceProblem:
    """A problem's operands on the GPU (values f64 for SpMV, f32 for SpMM)."""

    def __init__(self, A, op, K, seed=0):
        import torch
        from . import descriptors as D
        self.op, self.K, self.m, self.n, self.nnz = op, K, A.m, A.n, A.nnz
        dt = torch.float64 if op == 0 else torch.float32
        name = "f64" if op == 0 else "f32"
        pos = torch.from_numpy(A.pos).cuda()
        crd = torch.from_numpy(A.crd).cuda()
        vals = torch.from_numpy(A.vals).to(dt).cuda()
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        self.a = D.csr(pos, crd, vals, A.shape, name)
        if op == 0:
            self.inp = torch.rand(A.n, dtype=dt, device="cuda", generator=g) * 2 - 1
            self.out = torch.empty(A.m, dtype=dt, device="cuda")
            self.bd, self.cd = D.dense(self.inp, name), D.dense(self.out, name)
        else:
            self.inp = torch.rand((A.n, K), dtype=dt, device="cuda", generator=g) * 2 - 1
            self.out = torch.empty((A.m, K), dtype=dt, device="cuda")
            self.bd, self.cd = D.dense(self.inp, name), D.dense(self.out, name)
        self._keep = (pos, crd, vals)

    def runner(self, c):
        """The call that runs class c (None: the slab plan does not apply to
        this matrix).  The slab plan is built on first use; its build time
        is self.plan_build_ms."""
        L = _lib.lib()
        s = class_schedule(c, self.K)
        if self.op == 0:
            if s is None and not hasattr(self, "plan"):
                self.plan, self.plan_build_ms = autotune.slab_plan(self.a)
            return autotune.spmv_runner(s, self.a, self.bd, self.cd,
                                        getattr(self, "plan", None))
        return lambda: _lib.check(L.taco_hip_spmm_compute(self.cd.ptr, self.a.ptr, self.bd.ptr, s))


def time_classes(prob, reps=3, slow_factor=4.0, rtol=None):
    """Exhaustive timing of the problem's classes (ms; inf = rejected).

    The op's first class (one thread per row: the sequential summation
    order) runs once to give the reference output.  Every other class runs
    once and its output is checked against the reference to `rtol` relative
    to the largest output magnitude (a class that fails is rejected), then
    `reps` launches are timed between HIP events; a class slower than
    slow_factor x the best so far after its first timed launch keeps that
    single time.  The reference class is timed last, the same way."""
    import torch
    lo, hi = class_range(prob.op)
    rtol = rtol if rtol is not None else (1e-10 if prob.op == 0 else 1e-4)
    times = np.full(N_CLASSES, np.inf)
    e = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
    prob.runner(lo)()
    torch.cuda.synchronize()
    ref = prob.out.clone()
    scale = max(float(ref.abs().max().item()) if ref.numel() else 0.0, 1e-30)
    best = np.inf
    for c in list(range(lo + 1, hi)) + [lo]:
        fn = prob.runner(c)
        if fn is None:
            continue
        if c != lo:
            fn()
            torch.cuda.synchronize()
            diff = float((prob.out - ref).abs().max().item()) if ref.numel() else 0.0
            if diff > rtol * scale:
                continue
        e[0].record()
        fn()
        e[1].record()
        torch.cuda.synchronize()
        t1 = e[0].elapsed_time(e[1])
        if t1 > slow_factor * best or reps == 1:
            times[c] = t1
        else:
            for r in range(1, reps):
                fn()
                e[r + 1].record()
            torch.cuda.synchronize()
            ts = [t1] + [e[r].elapsed_time(e[r + 1]) for r in range(1, reps)]
            times[c] = float(np.median(ts))
        if c == PLAN_CLASS and prob.op == 0:
            times[c] += prob.plan_build_ms / autotune.PLAN_CALLS
        best = min(best, times[c])
    return times


# ------------------------------------------------------------ training --

def standardize(F, mu, sigma):
    return ((F - mu) / sigma).astype(np.float32)


def train(F, times, epochs=300, batch=256, lr=2e-3, weight_decay=1e-5, seed=0,
          log=None):
    """Trains on features F[N, 32] and class times[N, 24] (ms, inf =
    unavailable).  Returns (params, mu, sigma, loss history)."""
    import torch
    F = np.asarray(F, dtype=np.float64)
    mu = F.mean(axis=0)
    sigma = F.std(axis=0)
    sigma[sigma < 1e-6] = 1.0
    X = torch.from_numpy(standardize(F, mu, sigma)).cuda()
    best = np.argmin(times, axis=1).astype(np.int32)
    logt = np.log(np.min(times, axis=1)).astype(np.float32)
    cls = torch.from_numpy(best).cuda()
    lt = torch.from_numpy(logt).cuda()
    net = Mlp(seed=seed)
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    hist = []
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(n)).cuda()
        Xs, cs, ls = X[perm].contiguous(), cls[perm].contiguous(), lt[perm].contiguous()
        rate = lr * 0.5 * (1 + np.cos(np.pi * ep / epochs))
        tot = 0.0
        nb = 0
        for b0 in range(0, n, batch):
            b1 = min(n, b0 + batch)
            want = ep % 10 == 0 or ep == epochs - 1
            loss = net.train_step(Xs[b0:b1], cs[b0:b1], ls[b0:b1], rate, weight_decay,
                                  want_loss=want)
            if want:
                tot += loss * (b1 - b0)
                nb += b1 - b0
        if nb:
            hist.append((ep, tot / nb))
            if log:
                log(f"epoch {ep}: loss {tot / nb:.4f}")
    return net.params(), mu, sigma, hist


def save(path, params, mu, sigma, meta=None):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=params, mu=mu, sigma=sigma, classes=np.array(CLASS_NAMES),
             meta=np.array(repr(meta or {})))


class Selector:
    """Loads trained weights and picks a schedule for an operand."""

    def __init__(self, path=DEFAULT_WEIGHTS):
        d = np.load(path, allow_pickle=False)
        if list(d["classes"]) != CLASS_NAMES:
            raise ValueError(f"{path}: class list does not match this library")
        self.mu, self.sigma = d["mu"], d["sigma"]
        self.net = Mlp(d["params"])

    def scores(self, F):
        """Network outputs [N, 25] for raw features F[N, 32]."""
        import torch
        X = torch.from_numpy(standardize(np.atleast_2d(F), self.mu, self.sigma)).cuda()
        return self.net.forward(X).cpu().numpy()

    def choose_from_features(self, F):
        Z = self.scores(F)
        F = np.atleast_2d(F)
        out = np.empty(len(F), dtype=np.int64)
        for i in range(len(F)):
            lo, hi = class_range(int(round(F[i, 0])))
            out[i] = lo + int(np.argmax(Z[i, lo:hi]))
        return out, Z[:, N_CLASSES]

    def choose(self, a_desc, op, K=0):
        """(class, schedule, predicted ms) for one operand; bf16 SpMM takes
        the MFMA rule (no prediction: nan)."""
        if mfma_rule(a_desc, op, K):
            return RULE_MFMA, class_schedule(RULE_MFMA, K), float("nan")
        f = features(a_desc, op, K)
        c, lt = self.choose_from_features(f)
        c = int(c[0])
        return c, class_schedule(c, K), float(np.exp(lt[0]))


def evaluate(choice, times, ops):
    """Totals over problems: learned, exhaustive-best and default times (ms),
    plus per-problem ratios."""
    idx = np.arange(len(times))
    learned = times[idx, choice]
    best = times.min(axis=1)
    default = np.array([times[i, DEFAULT_CLASS[int(o)]] for i, o in enumerate(ops)])
    return {
        "learned_ms": float(learned.sum()), "exhaustive_ms": float(best.sum()),
        "default_ms": float(default.sum()),
        "learned_over_exhaustive": float(learned.sum() / best.sum()),
        "default_over_learned": float(default.sum() / learned.sum()),
        "geomean_learned_over_exhaustive": float(np.exp(np.mean(np.log(learned / best)))),
        "geomean_default_over_learned": float(np.exp(np.mean(np.log(default / learned)))),
        "top1": float(np.mean(choice == times.argmin(axis=1))),
    }
