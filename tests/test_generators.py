"""Seeded generators: structure, determinism and the workload statistics."""
import numpy as np

from taco_amd import workloads as W


def test_fixed_rows_distinct_sorted_in_range():
    A = W.csr_fixed(5000, 300, 10, 3)
    assert A.nnz == 50000
    for i in range(0, 5000, 97):
        r = A.crd[A.pos[i]:A.pos[i + 1]]
        assert len(r) == 10 and np.all(np.diff(r) > 0) and r.min() >= 0 and r.max() < 300
    assert np.all((A.vals >= -1) & (A.vals < 1))


def test_dense_rows_and_complement_path():
    A = W.csr_from_counts(50, 20, np.array([20] * 25 + [15] * 25, np.int32), 4)
    for i in range(50):
        r = A.crd[A.pos[i]:A.pos[i + 1]]
        assert np.all(np.diff(r) > 0)
    assert np.array_equal(A.crd[:20], np.arange(20))


def test_deterministic():
    a = W.csr_powerlaw(20000, 20000, seed=5)
    b = W.csr_powerlaw(20000, 20000, seed=5)
    c = W.csr_powerlaw(20000, 20000, seed=6)
    assert np.array_equal(a.crd, b.crd) and np.array_equal(a.vals, b.vals)
    assert not np.array_equal(a.pos, c.pos)


def test_powerlaw_statistics():
    counts = W.zipf_counts(400000, 2.0, 1, 100000, 16.0, 9)
    assert counts.min() >= 1 and counts.max() <= 100000
    assert abs(counts.mean() - 16.0) < 1.5
    # heavy tail: a visible share of entries in rows longer than 1000
    assert counts[counts > 1000].sum() / counts.sum() > 0.1


def test_banded_and_block_structure():
    counts = W.uniform_counts(3000, 2, 64, 7)
    A = W.csr_from_counts(3000, 3000, counts, 7, W.STRUCT_BANDED, 64)
    rows = np.repeat(np.arange(3000), np.diff(A.pos))
    assert np.all(np.abs(A.crd - rows) <= 64)
    B = W.csr_from_counts(3000, 3000, np.full(3000, 8, np.int32), 8, W.STRUCT_BLOCK, 8)
    rows = np.repeat(np.arange(3000), 8)
    assert np.array_equal(B.crd, (rows // 8) * 8 + np.tile(np.arange(8), 3000))
