"""Selector kernels against their CPU restatement, training determinism,
and the trained selector: unchanged results and its gain."""
import os

import numpy as np
import pytest
import torch

import oracle
from oracle import selector_ref as R
from taco_amd import autotune
from taco_amd import _lib, selector as S
from taco_amd import descriptors as D
from taco_amd import workloads as W
from tests import gpu_util as G

pytestmark = pytest.mark.gpu


def dev_csr(A):
    return G.csr_desc(A)


@pytest.mark.parametrize("seed", range(8))
def test_stats_and_features_match_restatement(seed):
    A, op, K, d = S.make_problem(seed, 3, rmin=3000, rmax=60000)
    print(f"stats seed={seed} {d}", flush=True)
    a = dev_csr(A)
    s_gpu = S.csr_stats(a)
    s_ref = R.csr_stats(A.pos, A.crd, A.m, A.n)
    assert np.array_equal(s_gpu, s_ref), (s_gpu, s_ref)
    np.testing.assert_allclose(S.features(a, op, K), R.features_from_stats(s_ref, op, K),
                               rtol=1e-6, atol=1e-7)


def test_stats_edge_cases():
    cases = [W.Csr(5, 7, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0)),
             W.csr_from_counts(1, 100000, np.array([100000], np.int32), 3),
             W.csr_from_counts(9000, 50, W.uniform_counts(9000, 0, 50, 4), 4)]
    for A in cases:
        s_gpu = S.csr_stats(dev_csr(A))
        assert np.array_equal(s_gpu, R.csr_stats(A.pos, A.crd, A.m, A.n))


def batch(seed, B):
    rng = np.random.default_rng(seed)
    p = S.init_params(seed)
    p = (p + rng.normal(0, 0.05, p.size)).astype(np.float32)
    X = rng.normal(0, 1, (B, 32)).astype(np.float32)
    cls = np.concatenate([rng.integers(0, 13, B // 2), rng.integers(13, 24, B - B // 2)])
    logt = rng.normal(-1, 1, B).astype(np.float32)
    return p, X, cls.astype(np.int32), logt


def test_forward_matches_cpu():
    p, X, _, _ = batch(1, 300)
    net = S.Mlp(p)
    Z = net.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    Zr = R.forward(p, X)[3]
    assert np.max(np.abs(Z - Zr)) <= 1e-4 * max(1.0, np.max(np.abs(Zr)))


@pytest.mark.parametrize("B", [1, 37, 256])
def test_gradients_match_cpu(B):
    p, X, cls, logt = batch(2 + B, B)
    net = S.Mlp(p)
    loss, g = net.gradients(torch.from_numpy(X).cuda(), torch.from_numpy(cls).cuda(),
                            torch.from_numpy(logt).cuda())
    loss_r, g_r = R.loss_and_grad(p, X, cls, logt)
    assert abs(loss - loss_r) <= 1e-5 * max(1.0, abs(loss_r))
    # per parameter block, relative to the block's largest gradient
    at = 0
    for size in (4096, 128, 16384, 128, 3200, 25):
        blk = slice(at, at + size)
        scale = max(np.max(np.abs(g_r[blk])), 1e-12)
        assert np.max(np.abs(g[blk] - g_r[blk])) <= 1e-4 * scale, (at, size)
        at += size


def synthetic_db(n, seed):
    """Features whose best class is a simple function of two features."""
    rng = np.random.default_rng(seed)
    F = rng.normal(0, 1, (n, 32)).astype(np.float32)
    F[:, 0] = rng.integers(0, 2, n)
    T = np.full((n, 24), np.inf)
    for i in range(n):
        lo, hi = S.class_range(int(F[i, 0]))
        T[i, lo:hi] = 2.0
        k = lo + (int(F[i, 5] > 0) + 2 * int(F[i, 7] > 0)) % (hi - lo)
        T[i, k] = 1.0
    return F, T


def test_training_is_deterministic_and_learns():
    F, T = synthetic_db(600, 5)
    p1, _, _, h1 = S.train(F, T, epochs=40, batch=64, lr=3e-3, seed=3)
    p2, _, _, h2 = S.train(F, T, epochs=40, batch=64, lr=3e-3, seed=3)
    assert np.array_equal(p1, p2), "training is not bitwise reproducible"
    assert h1[-1][1] < 0.25 * h1[0][1], h1
    net = S.Mlp(p1)
    mu, sigma = F.astype(np.float64).mean(0), F.astype(np.float64).std(0)
    Z = net.forward(torch.from_numpy(S.standardize(F, mu, sigma)).cuda()).cpu().numpy()
    choice = np.array([S.class_range(int(f[0]))[0] + np.argmax(
        z[S.class_range(int(f[0]))[0]:S.class_range(int(f[0]))[1]]) for f, z in zip(F, Z)])
    assert np.mean(choice == T.argmin(1)) > 0.95


needs_weights = pytest.mark.skipif(not os.path.exists(S.DEFAULT_WEIGHTS),
                                   reason="no trained selector weights in taco_amd/data")


@needs_weights
def test_selected_schedule_changes_no_result():
    sel = S.Selector()
    for i in range(6):
        A, op, K, d = S.make_problem(99, i, rmin=5000, rmax=50000)
        prob = S.DeviceProblem(A, op, K, seed=i)
        c, sched, _ = sel.choose(prob.a, op, K)
        print(f"selected {S.CLASS_NAMES[c]} for {d}", flush=True)
        prob.runner(c)()
        torch.cuda.synchronize()
        got = prob.out.cpu().numpy().astype(np.float64)
        prob.out.fill_(float("nan"))
        L = _lib.lib()
        if sched is None:  # the slab plan, built again
            plan, _ = autotune.slab_plan(prob.a)
            plan.compute(prob.cd, prob.bd)
        else:
            fn = L.taco_hip_spmv_compute if op == 0 else L.taco_hip_spmm_compute
            _lib.check(fn(prob.cd.ptr, prob.a.ptr, prob.bd.ptr, sched))
        torch.cuda.synchronize()
        again = prob.out.cpu().numpy().astype(np.float64)
        if c not in S.ATOMIC_CLASSES:  # fixed summation order: bit for bit
            assert np.array_equal(got, again)
        Ah = W.Csr(A.m, A.n, A.pos, A.crd, A.vals if op == 0 else
                   A.vals.astype(np.float32).astype(np.float64))
        inp = prob.inp.cpu().numpy().astype(np.float64)
        for out in (got, again):
            if op == 0:
                ye, cond = oracle.exact_spmv(Ah, inp)
                rep = oracle.error_report(out, ye, cond, 1e-12)
            else:
                ye, cond = oracle.exact_spmm(Ah, inp)
                rep = oracle.error_report(out.reshape(A.m, K), ye, cond, 1e-5)
            assert rep["ok"], (d, S.CLASS_NAMES[c], rep)


@needs_weights
def test_selector_gain_over_default():
    """On fresh problems the learned choice beats the default schedule and
    stays near the exhaustive search (loose bounds: short timings)."""
    sel = S.Selector()
    F, T, ops = [], [], []
    for i in range(16):
        A, op, K, d = S.make_problem(123, i, rmin=50_000, rmax=400_000)
        prob = S.DeviceProblem(A, op, K, seed=i)
        F.append(S.features(prob.a, op, K))
        T.append(S.time_classes(prob, reps=5))
        ops.append(op)
    choice, _ = sel.choose_from_features(np.array(F))
    ev = S.evaluate(choice, np.array(T), np.array(ops))
    print(ev, flush=True)
    assert ev["default_over_learned"] > 1.2
    assert ev["learned_over_exhaustive"] < 1.25
