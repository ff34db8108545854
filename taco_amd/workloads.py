"""Seeded synthetic workloads of the sparse hot path (host numpy arrays).

All randomness comes from libtaco_hip's counter-based generators, so a given
(seed, sizes) always produces the same matrices on any machine and thread
count.  Streams keep the independent draws apart.
"""
import ctypes

import numpy as np

from . import _lib

STREAM_VALS = 11
STREAM_X = 12
STREAM_B = 13
STREAM_VALS_B = 14

STRUCT_UNIFORM, STRUCT_BANDED, STRUCT_BLOCK = 0, 1, 2


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def uniform(n, seed, stream, lo=-1.0, hi=1.0):
    out = np.empty(int(n), dtype=np.float64)
    if n:
        _lib.lib().taco_hip_gen_uniform_f64(int(n), seed, stream, lo, hi, _p(out))
    return out


class Csr:
    """Host CSR matrix: pos int32[m+1], crd int32[nnz], vals float64[nnz]."""

    def __init__(self, m, n, pos, crd, vals):
        self.m, self.n = int(m), int(n)
        self.pos, self.crd, self.vals = pos, crd, vals

    @property
    def nnz(self):
        return int(self.pos[-1])

    @property
    def shape(self):
        return (self.m, self.n)

    def dense(self):
        d = np.zeros((self.m, self.n), dtype=self.vals.dtype)
        rows = np.repeat(np.arange(self.m), np.diff(self.pos))
        d[rows, self.crd] = self.vals
        return d


def counts_to_pos(counts):
    pos = np.empty(len(counts) + 1, dtype=np.int32)
    if _lib.lib().taco_hip_counts_to_pos(len(counts), _p(counts), _p(pos)) != 0:
        raise ValueError("entry count exceeds the int32 index range")
    return pos


def csr_from_counts(m, n, counts, seed, structure=STRUCT_UNIFORM, band=0,
                    values=True):
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    pos = counts_to_pos(counts)
    crd = np.empty(int(pos[-1]), dtype=np.int32)
    rc = _lib.lib().taco_hip_gen_csr_columns(m, n, _p(pos), structure, band,
                                             seed, _p(crd))
    if rc != 0:
        raise ValueError("a row asks for more columns than its structure holds")
    vals = uniform(len(crd), seed, STREAM_VALS) if values else None
    return Csr(m, n, pos, crd, vals)


def csr_fixed(m, n, k, seed):
    """Exactly k distinct uniform columns per row (sorted), values U[-1,1]."""
    return csr_from_counts(m, n, np.full(m, k, dtype=np.int32), seed)


def zipf_counts(m, alpha, kmin, kmax, mean, seed):
    counts = np.empty(m, dtype=np.int32)
    total = _lib.lib().taco_hip_gen_row_counts_zipf(m, alpha, kmin, kmax, mean,
                                                    seed, _p(counts))
    if total < 0:
        raise ValueError("bad Zipf parameters")
    return counts


def csr_powerlaw(m, n, seed, alpha=2.0, kmin=1, kmax=100000, mean=16.0):
    """Row lengths Zipf(alpha) on [kmin, kmax] rescaled to `mean`; uniform
    distinct sorted columns; values U[-1,1]."""
    counts = zipf_counts(m, alpha, kmin, min(kmax, n), mean, seed)
    return csr_from_counts(m, n, counts, seed)


def uniform_counts(m, kmin, kmax, seed):
    counts = np.empty(m, dtype=np.int32)
    _lib.lib().taco_hip_gen_row_counts_uniform(m, kmin, kmax, seed, _p(counts))
    return counts


# The workloads of the hot path, by name.
def spmv_uniform_100k(seed=1):
    """100,000 x 100,000, exactly 10 nnz/row (1M nnz); x U[-1,1]."""
    A = csr_fixed(100_000, 100_000, 10, seed)
    x = uniform(A.n, seed, STREAM_X)
    return A, x


def powerlaw_2m(seed=2):
    """2,000,000 x 2,000,000, Zipf(2.0) rows on [1, 1e5] rescaled to mean 16."""
    return csr_powerlaw(2_000_000, 2_000_000, seed)


def spgemm_500k(seed=3):
    """A, B 500,000 x 500,000 with 8 uniform distinct columns per row."""
    A = csr_fixed(500_000, 500_000, 8, seed)
    B = csr_fixed(500_000, 500_000, 8, seed + 1000)
    B.vals = uniform(B.nnz, seed, STREAM_VALS_B)
    return A, B


def algorithmic_bytes_spmv(m, n, nnz, val_bytes=8):
    """Compulsory traffic: pos (m+1) int32, crd/vals once, x and y once."""
    return 4 * (m + 1) + (4 + val_bytes) * nnz + val_bytes * n + val_bytes * m


def algorithmic_bytes_spmm(m, n, nnz, k, a_bytes=8, b_bytes=8, c_bytes=8):
    return 4 * (m + 1) + (4 + a_bytes) * nnz + b_bytes * n * k + c_bytes * m * k
