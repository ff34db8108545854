"""Pins the CPU oracle (oracle/taco_cpu.c, oracle/oracle_exact.c).

Each oracle function is checked against (a) the reference's known-answer
tests, (b) the golden fixtures written by the real taco (tests/golden), (c)
the real taco itself where oracle/_ref was built, and (d) what mathematics
fixes: closed forms, brute force on dense matrices, invariants.  A planted
mistake in an oracle result must fail the gate.
"""
import os

import numpy as np
import pytest

import oracle
from taco_amd import workloads as W
from tests import known_answers as KA

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def csr(m, n, entries):
    p, c, v = KA.coo_to_csr(m, n, entries)
    return W.Csr(m, n, p, c, v)


def dense_vec(n, d):
    x = np.zeros(n)
    for k, v in d.items():
        x[k] = v
    return x


@pytest.fixture(scope="module", autouse=True)
def threads():
    oracle.set_threads(4)


# ---- (a) known answers from the reference's tests ----

def test_known_spmv_d33a_d3b():
    A = csr(3, 3, KA.D33A)
    y = oracle.spmv(A, dense_vec(3, KA.D3B))
    assert y.tolist() == [0.0, 0.0, 18.0]


def test_known_matrix_vector_mul():
    A = csr(5, 5, KA.MVM_B)
    y = oracle.spmv(A, dense_vec(5, KA.MVM_C))
    assert y.tolist() == [2.0, 0.0, 0.0, 8.0, 0.0]


def test_known_matrix_mul():
    A = csr(3, 3, KA.D33A)
    B = np.zeros((3, 3))
    for (i, j), v in KA.D33B:
        B[i, j] = v
    C = oracle.spmm(A, B)
    assert C.tolist() == [[0, 0, 0], [0, 0, 0], [30, 180, 0]]


def test_known_where_spmm_spgemm():
    A = csr(5, 5, KA.SPMM_B)
    B = csr(5, 5, KA.SPMM_C)
    pos, crd = oracle.spgemm_assemble(A, B)
    vals = oracle.spgemm_compute(A, B, pos, crd)
    assert pos.tolist() == KA.SPMM_EXPECTED_POS
    assert crd.tolist() == KA.SPMM_EXPECTED_CRD
    assert vals.tolist() == [30.0, 180.0]


# ---- (b) golden fixtures written by the real taco ----

def test_golden_reference_test_inputs():
    d = np.load(os.path.join(GOLD, "ref_tests.npz"))
    for tag in ("spmvgpu", "spmvcpu"):
        A = W.Csr(102, 103, d[f"{tag}_pos"], d[f"{tag}_crd"], d[f"{tag}_vals"])
        y = oracle.spmv(A, d[f"{tag}_x"])
        assert np.array_equal(y, d[f"{tag}_y"]), tag
    A = W.Csr(102, 103, d["spmmgpu_pos"], d["spmmgpu_crd"], d["spmmgpu_vals"])
    C = oracle.spmm(A, d["spmmgpu_B"])
    assert np.array_equal(C, d["spmmgpu_C"])


def test_golden_seeded_bit_exact():
    """Same loop nest and summation order as taco's generated C: bit-equal."""
    d = np.load(os.path.join(GOLD, "seeded.npz"))
    A = W.Csr(len(d["spmv_pos"]) - 1, len(d["spmv_x"]), d["spmv_pos"],
              d["spmv_crd"], d["spmv_vals"])
    assert np.array_equal(oracle.spmv(A, d["spmv_x"]), d["spmv_y"])
    assert np.array_equal(oracle.spmm(A, d["spmm_B"]), d["spmm_C"])
    G1 = W.Csr(len(d["g_apos"]) - 1, len(d["g_bpos"]) - 1, d["g_apos"],
               d["g_acrd"], d["g_avals"])
    G2 = W.Csr(len(d["g_bpos"]) - 1, int(d["g_bn"][0]), d["g_bpos"],
               d["g_bcrd"], d["g_bvals"])
    pos, crd = oracle.spgemm_assemble(G1, G2)
    assert np.array_equal(pos, d["g_cpos"])
    assert np.array_equal(crd, d["g_ccrd"])
    assert np.array_equal(oracle.spgemm_compute(G1, G2, pos, crd), d["g_cvals"])


# ---- (c) the real taco, where it was built here ----

needs_ref = pytest.mark.skipif(not oracle.ref_available(),
                               reason="oracle/_ref not built (no reference tree)")


@needs_ref
@pytest.mark.parametrize("seed", [5, 6])
def test_ref_spmv_spmm_bit_exact(seed):
    A = W.csr_powerlaw(4000, 3000, seed=seed, kmax=2500)
    x = W.uniform(A.n, seed, W.STREAM_X)
    assert np.array_equal(oracle.spmv(A, x), oracle.ref_spmv(A, x, threads=3))
    B = W.uniform(A.n * 8, seed, W.STREAM_B).reshape(A.n, 8)
    assert np.array_equal(oracle.spmm(A, B), oracle.ref_spmm(A, B, threads=3))


@needs_ref
def test_ref_spgemm_bit_exact():
    A = W.csr_powerlaw(2000, 1500, seed=9, kmax=200)
    B = W.csr_fixed(1500, 1700, 5, 10)
    pos, crd = oracle.spgemm_assemble(A, B)
    vals = oracle.spgemm_compute(A, B, pos, crd)
    rp, rc, rv = oracle.ref_spgemm(A, B, threads=3)
    assert np.array_equal(pos, rp) and np.array_equal(crd, rc)
    assert np.array_equal(vals, rv)


# ---- (d) closed forms, brute force, invariants ----

def test_spmv_ones_gives_row_sums():
    A = W.csr_powerlaw(3000, 3000, seed=3, kmax=1000)
    y = oracle.spmv(A, np.ones(A.n))
    rows = np.add.reduceat(A.vals, A.pos[:-1]) if A.nnz else np.zeros(A.m)
    rows[np.diff(A.pos) == 0] = 0
    np.testing.assert_allclose(y, rows, rtol=1e-13, atol=1e-13)


def test_spmv_permutation_gives_permuted_x():
    rng = np.random.default_rng(1)
    perm = rng.permutation(500).astype(np.int32)
    A = W.Csr(500, 500, np.arange(501, dtype=np.int32), perm, np.ones(500))
    x = W.uniform(500, 1, 2)
    assert np.array_equal(oracle.spmv(A, x), x[perm])


def test_spmv_spmm_match_dense_brute_force():
    A = W.csr_powerlaw(300, 250, seed=4, kmax=250)
    D = A.dense()
    x = W.uniform(250, 4, 1)
    np.testing.assert_allclose(oracle.spmv(A, x), D @ x, rtol=0, atol=1e-12)
    B = W.uniform(250 * 7, 4, 2).reshape(250, 7)
    np.testing.assert_allclose(oracle.spmm(A, B), D @ B, rtol=0, atol=1e-12)
    # SpMM with B = e_k columns reduces to SpMV with x = B[:, k]
    C = oracle.spmm(A, B)
    for k in range(7):
        assert np.array_equal(C[:, k], oracle.spmv(A, np.ascontiguousarray(B[:, k])))


def test_spgemm_identity_and_structure():
    A = W.csr_powerlaw(400, 300, seed=5, kmax=300)
    eye = W.Csr(300, 300, np.arange(301, dtype=np.int32),
                np.arange(300, dtype=np.int32), np.ones(300))
    pos, crd = oracle.spgemm_assemble(A, eye)
    assert np.array_equal(pos, A.pos) and np.array_equal(crd, A.crd)
    assert np.array_equal(oracle.spgemm_compute(A, eye, pos, crd), A.vals)
    # pattern row i = sorted union of B's rows named by A's row i
    B = W.csr_fixed(300, 350, 6, 6)
    pos, crd = oracle.spgemm_assemble(A, B)
    assert pos[0] == 0 and np.all(np.diff(pos) >= 0)
    for i in range(A.m):
        want = set()
        for p in range(A.pos[i], A.pos[i + 1]):
            j = A.crd[p]
            want.update(B.crd[B.pos[j]:B.pos[j + 1]].tolist())
        assert crd[pos[i]:pos[i + 1]].tolist() == sorted(want)
    vals = oracle.spgemm_compute(A, B, pos, crd)
    D = A.dense() @ B.dense()
    rows = np.repeat(np.arange(A.m), np.diff(pos))
    np.testing.assert_allclose(vals, D[rows, crd], rtol=0, atol=1e-12)


def test_spgemm_pattern_is_structural():
    """Cancelling products keep their entry (taco's Bool-dedup count query)."""
    A = W.Csr(1, 2, np.array([0, 2], np.int32), np.array([0, 1], np.int32),
              np.array([1.0, -1.0]))
    B = W.Csr(2, 1, np.array([0, 1, 2], np.int32), np.array([0, 0], np.int32),
              np.array([1.0, 1.0]))
    pos, crd = oracle.spgemm_assemble(A, B)
    assert pos.tolist() == [0, 1] and crd.tolist() == [0]
    assert oracle.spgemm_compute(A, B, pos, crd).tolist() == [0.0]


def test_exact_oracle_against_fractions():
    """The extended-precision oracle against exact rational arithmetic."""
    from fractions import Fraction
    A = W.csr_powerlaw(60, 80, seed=8, kmax=80)
    x = W.uniform(80, 8, 3)
    y, cond = oracle.exact_spmv(A, x)
    for i in range(A.m):
        s = sum((Fraction(A.vals[p]) * Fraction(x[A.crd[p]])
                 for p in range(A.pos[i], A.pos[i + 1])), Fraction(0))
        assert y[i] == float(s)
        c = sum(abs(float(Fraction(A.vals[p]) * Fraction(x[A.crd[p]])))
                for p in range(A.pos[i], A.pos[i + 1]))
        assert abs(cond[i] - c) <= 1e-15 * max(c, 1e-300)


def test_exact_spmm_against_fractions():
    """exact_spmm against exact rational arithmetic, every (i, k), on a
    seeded power-law matrix with empty, short and long rows."""
    from fractions import Fraction
    P = W.csr_powerlaw(50, 70, seed=21, kmax=70)
    # one empty row appended after short and long ones
    A = W.Csr(51, 70, np.append(P.pos, P.pos[-1]).astype(np.int32), P.crd, P.vals)
    assert np.diff(A.pos).max() > 20 and np.diff(A.pos)[-1] == 0
    K = 5
    B = W.uniform(70 * K, 21, W.STREAM_B).reshape(70, K)
    C, cond = oracle.exact_spmm(A, B)
    for i in range(A.m):
        for k in range(K):
            terms = [Fraction(A.vals[p]) * Fraction(B[A.crd[p], k])
                     for p in range(A.pos[i], A.pos[i + 1])]
            assert C[i, k] == float(sum(terms, Fraction(0))), (i, k)
            c = sum(abs(float(t)) for t in terms)
            assert abs(cond[i, k] - c) <= 1e-15 * max(c, 1e-300), (i, k)


def test_exact_spmm_columns_are_exact_spmv():
    """Column k of exact_spmm(A, B) is exact_spmv(A, B[:, k]), bit for bit,
    values and condition sums, on a 4,000-row power-law matrix."""
    A = W.csr_powerlaw(4000, 3500, seed=22, kmax=3000)
    K = 9
    B = W.uniform(A.n * K, 22, W.STREAM_B).reshape(A.n, K)
    C, cond = oracle.exact_spmm(A, B)
    for k in range(K):
        y, c = oracle.exact_spmv(A, np.ascontiguousarray(B[:, k]))
        assert np.array_equal(C[:, k], y), k
        assert np.array_equal(cond[:, k], c), k


def test_exact_spmm_integer_golden():
    """The reference's SpMM GPU test inputs (tests-scheduling-eval.cpp:
    1258-1307) hold small integers, so every correct summation order gives
    the golden written by the real taco exactly."""
    d = np.load(os.path.join(GOLD, "ref_tests.npz"))
    A = W.Csr(102, 103, d["spmmgpu_pos"], d["spmmgpu_crd"], d["spmmgpu_vals"])
    C, _ = oracle.exact_spmm(A, d["spmmgpu_B"])
    assert np.array_equal(C, d["spmmgpu_C"])


def test_exact_spmm_gate_rejects_transposed_b():
    """A result computed with B indexed as if column-major (B[k*n + j]
    instead of B[j*K + k]) fails the 1e-12 gate against exact_spmm."""
    A = W.csr_powerlaw(3000, 3000, seed=23, kmax=800)
    K = 16
    B = W.uniform(A.n * K, 23, W.STREAM_B).reshape(A.n, K)
    Ce, cond = oracle.exact_spmm(A, B)
    assert oracle.error_report(oracle.spmm(A, B), Ce, cond, 1e-12)["ok"]
    # B'[j, k] = flat(B)[k * n + j]: the transposed-index mistake.
    Bt = np.ascontiguousarray(B.ravel().reshape(K, A.n).T)
    assert not oracle.error_report(oracle.spmm(A, Bt), Ce, cond, 1e-12)["ok"]
    # and a dropped last column term, on one long row
    r = int(np.argmax(np.diff(A.pos)))
    Cd = oracle.spmm(A, B)
    p = A.pos[r + 1] - 1
    Cd[r, K - 1] -= A.vals[p] * B[A.crd[p], K - 1]
    assert not oracle.error_report(Cd, Ce, cond, 1e-12)["ok"]


def test_exact_spgemm_values_against_fractions():
    from fractions import Fraction
    A = W.csr_fixed(30, 40, 5, 11)
    B = W.csr_fixed(40, 50, 4, 12)
    pos, crd = oracle.spgemm_assemble(A, B)
    vals, cond = oracle.exact_spgemm_values(A, B, pos, crd)
    D = {}
    for i in range(A.m):
        for p in range(A.pos[i], A.pos[i + 1]):
            j = A.crd[p]
            for q in range(B.pos[j], B.pos[j + 1]):
                D[(i, B.crd[q])] = D.get((i, B.crd[q]), Fraction(0)) + \
                    Fraction(A.vals[p]) * Fraction(B.vals[q])
    for i in range(A.m):
        for p in range(pos[i], pos[i + 1]):
            assert vals[p] == float(D[(i, crd[p])])


def test_gate_catches_planted_mistakes():
    """A dropped term, a wrong sign and a transposed index all fail the
    1e-12 condition-scaled gate."""
    A = W.csr_powerlaw(2000, 2000, seed=12, kmax=500)
    x = W.uniform(A.n, 12, 3)
    y = oracle.spmv(A, x)
    ye, cond = oracle.exact_spmv(A, x)
    assert oracle.error_report(y, ye, cond, 1e-12)["ok"]
    r = 7
    while A.pos[r + 1] - A.pos[r] < 2:
        r += 1
    dropped = y.copy()
    dropped[r] -= A.vals[A.pos[r]] * x[A.crd[A.pos[r]]]
    assert not oracle.error_report(dropped, ye, cond, 1e-12)["ok"]
    flipped = y.copy()
    flipped[r] = -flipped[r]
    assert not oracle.error_report(flipped, ye, cond, 1e-12)["ok"]
    At = W.Csr(A.m, A.n, A.pos, A.crd[::-1].copy(), A.vals)
    assert not oracle.error_report(oracle.spmv(At, x), ye, cond, 1e-12)["ok"]


def test_threads_do_not_change_results():
    A = W.csr_powerlaw(5000, 5000, seed=13, kmax=3000)
    x = W.uniform(A.n, 13, 3)
    oracle.set_threads(1)
    y1 = oracle.spmv(A, x)
    oracle.set_threads(7)
    y7 = oracle.spmv(A, x)
    oracle.set_threads(4)
    assert np.array_equal(y1, y7)
