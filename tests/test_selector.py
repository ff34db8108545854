"""Selector host logic and its CPU restatement (no GPU).

Pins the restatement of the MLP's backward pass against central finite
differences, the restated statistics against values worked out by hand, and
the library's host feature arithmetic against the restatement."""
import numpy as np
import pytest

from oracle import selector_ref as R
from taco_amd import selector as S
from taco_amd import workloads as W


def small_problem(seed=0, B=12):
    rng = np.random.default_rng(seed)
    p = S.init_params(seed).astype(np.float64)
    p += rng.normal(0, 0.05, p.size)  # non-zero biases too
    X = rng.normal(0, 1, (B, 32))
    cls = np.concatenate([rng.integers(0, 13, B // 2), rng.integers(13, 24, B - B // 2)])
    logt = rng.normal(-1, 1, B)
    return p, X, cls, logt


def test_backward_matches_finite_differences():
    p, X, cls, logt = small_problem()
    rng = np.random.default_rng(1)
    idx = rng.choice(p.size, 300, replace=False)
    # every block: W1, b1, W2, b2, W3 (class and time rows), b3
    idx = np.concatenate([idx, [0, 4096, 4100, 4224, 4352, 20736, 20737, 23935, 23960]])
    assert R.check_gradients(p, X, cls, logt, idx) < 1e-5


def test_planted_backward_mistake_is_caught():
    """A dropped ReLU mask in the restatement fails the check."""
    p, X, cls, logt = small_problem(2)
    orig = R.loss_and_grad

    def broken(p, X, cls, logt):
        loss, g = orig(p, X, cls, logt)
        W1, b1, W2, b2, W3, b3 = R.unpack(p)
        _, h1, h2, Z = R.forward(p, X)
        # recompute dW1 without the layer-1 ReLU mask
        B = X.shape[0]
        dZ = np.zeros_like(Z)
        for b in range(B):
            c = int(cls[b])
            lo, hi = (0, 13) if c < 13 else (13, 24)
            z = Z[b, lo:hi]
            pr = np.exp(z - z.max())
            pr /= pr.sum()
            pr[c - lo] -= 1
            dZ[b, lo:hi] = pr / B
            dZ[b, 24] = 2 * R.TIME_WEIGHT * (Z[b, 24] - logt[b]) / B
        d2 = (dZ @ W3) * (h2 > 0)
        d1 = d2 @ W2
        g = g.copy()
        g[:4096] = (d1.T @ X).ravel()
        return loss, g

    R.loss_and_grad = broken
    try:
        idx = np.arange(0, 4096, 97)
        assert R.check_gradients(p, X, cls, logt, idx) > 1e-3
    finally:
        R.loss_and_grad = orig


def test_loss_closed_form():
    """Zero weights: every logit and the time output is 0, so the loss is
    log(#classes of the op) + w * logt^2."""
    p = np.zeros(23961)
    X = np.ones((2, 32))
    loss, g = R.loss_and_grad(p, X, np.array([3, 20]), np.array([1.0, -2.0]))
    expect = (np.log(13) + 0.1 * 1.0 + np.log(11) + 0.1 * 4.0) / 2
    assert abs(loss - expect) < 1e-12


def test_stats_by_hand():
    # 4 x 10: rows of lengths 0, 1, 3, 70 is impossible (n=10); use n=100
    counts = np.array([0, 1, 3, 70], np.int32)
    A = W.csr_from_counts(4, 100, counts, 5)
    s = R.csr_stats(A.pos, A.crd, A.m, A.n)
    assert list(s[:7]) == [4, 100, 74, 74, 0 + 1 + 9 + 4900, 70, 1]
    hist = s[7:23]
    assert hist[0] == 1 and hist[1] == 1 and hist[2] == 1 and hist[7] == 1 and hist.sum() == 4
    assert s[23] == 70 and s[24] == 0
    assert s[25] == 4 and s[26] == 74
    # row r's diagonal column is r*100//4 = 25r; band / block counts by brute force
    rows = np.repeat(np.arange(4), counts)
    assert s[27] == np.count_nonzero(np.abs(A.crd - rows * 25) <= 64)
    assert s[28] == np.count_nonzero(A.crd // 8 == rows // 8)
    assert s[29] == np.sum(np.abs(A.crd.astype(np.int64) * 4 - rows * 100) * 4096 // 400)


def test_host_features_match_restatement():
    for seed in range(6):
        A, op, K, _ = S.make_problem(seed, 0, rmin=2000, rmax=20000)
        s = R.csr_stats(A.pos, A.crd, A.m, A.n)
        f_lib = S.features_from_stats(s, op, K)
        f_ref = R.features_from_stats(s, op, K)
        np.testing.assert_allclose(f_lib, f_ref, rtol=1e-6, atol=1e-7)


def test_problems_are_seeded_and_cover_structures():
    seen = set()
    for i in range(24):
        A, op, K, d = S.make_problem(11, i, rmin=1000, rmax=5000)
        B, op2, K2, d2 = S.make_problem(11, i, rmin=1000, rmax=5000)
        assert d == d2 and np.array_equal(A.crd, B.crd)
        assert (op == 0) == (K == 0)
        seen.add(d["structure"])
    assert seen == set(S.STRUCTURES)


def test_problem_op_matches_generated_problem():
    for i in range(40):
        assert S.problem_op(7, i) == S.make_problem(7, i, rmin=1000, rmax=3000)[1]


def test_class_space():
    assert len(S.CLASS_NAMES) == 24
    assert S.DEFAULT_CLASS[0] == 0 and S.DEFAULT_CLASS[1] == 13
    for K in S.SPMM_WIDTHS:
        for c in range(13, 24):
            s = S.class_schedule(c, K)
            assert s.group_size == 0 or s.group_size * 4 >= K


def test_bf16_spmm_rule():
    """bf16 SpMM with K a multiple of 4 goes to the MFMA kernel by rule;
    every other operand is left to the network."""
    import torch
    from taco_amd import _lib, descriptors as D
    pos = torch.zeros(3, dtype=torch.int32)
    crd = torch.zeros(0, dtype=torch.int32)
    for dtype, vals in (("bf16", torch.zeros(0, dtype=torch.int16)),
                        ("f32", torch.zeros(0, dtype=torch.float32)),
                        ("f64", torch.zeros(0, dtype=torch.float64))):
        a = D.csr(pos, crd, vals, (2, 2), dtype)
        assert S.mfma_rule(a, 1, 64) == (dtype == "bf16")
        assert not S.mfma_rule(a, 0, 0)
        if dtype == "bf16":
            assert not S.mfma_rule(a, 1, 30)
            c, sched, pred = S.Selector.choose(None, a, 1, 64)
            assert c == S.RULE_MFMA and sched.algo == _lib.MFMA and pred != pred
            assert S.class_name(c).startswith("spmm_mfma_bf16")
    assert S.class_name(0) == S.CLASS_NAMES[0]


def test_evaluate_totals():
    t = np.full((3, 24), np.inf)
    t[0, :13] = np.arange(13) + 1.0      # default (class 0) = 1, best = 1
    t[1, :13] = 10 - np.arange(13) * 0.5  # default 10, best 4 at class 12
    t[2, 13:] = np.arange(11) + 2.0      # default (13) = 2
    ev = S.evaluate(np.array([0, 12, 14]), t, np.array([0, 0, 1]))
    assert ev["learned_ms"] == 1 + 4 + 3 and ev["exhaustive_ms"] == 1 + 4 + 2
    assert ev["default_ms"] == 1 + 10 + 2
    assert ev["top1"] == pytest.approx(2 / 3)
