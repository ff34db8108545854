"""Unified-memory mode on the GPU (reference src/cuda.cpp:11-72,
src/storage/array.cpp:212-219): with unified memory enabled, taco allocates
every tensor array with cuda_unified_alloc and writes it from the host
(pack), then hands the pointers to generated code.  Here every operand and
output is allocated with taco_hip_unified_alloc (managed memory), filled and
read back by plain host copies, and the kernels that add into their output
with float atomics (the nnz-split and MFMA schedules, generated nnz-split and
merge modules) run on it.  Results are checked against the FP64 oracle."""
import ctypes

import numpy as np
import pytest

import oracle
from taco_amd import _lib, codegen, descriptors as D, workloads as W
from tests import gpu_util as G

pytestmark = pytest.mark.gpu


class Managed:
    """A managed-memory array written and read by host copies."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.dtype, self.shape, self.nbytes = host.dtype, host.shape, host.nbytes
        self.addr = _lib.lib().taco_hip_unified_alloc(max(self.nbytes, 8))
        assert self.addr, _lib.lib().taco_hip_last_error()
        ctypes.memmove(self.addr, host.ctypes.data, self.nbytes)

    def read(self):
        _lib.check(_lib.lib().taco_hip_synchronize())
        out = np.empty(self.shape, self.dtype)
        ctypes.memmove(out.ctypes.data, self.addr, self.nbytes)
        return out

    def free(self):
        _lib.lib().taco_hip_unified_free(ctypes.c_void_p(self.addr))
        self.addr = None


@pytest.fixture
def unified():
    L = _lib.lib()
    L.taco_hip_set_unified_memory_enabled(1)
    held = []
    try:
        yield held
    finally:
        for a in held:
            a.free()
        L.taco_hip_set_unified_memory_enabled(0)


def managed(held, host):
    a = Managed(host)
    held.append(a)
    return a


def csr_managed(held, A, dtype="f64"):
    if dtype == "bf16":
        vals = G.bf16_bits(A.vals).numpy()
    else:
        vals = A.vals.astype(np.float64 if dtype == "f64" else np.float32)
    p, c, v = managed(held, A.pos), managed(held, A.crd), managed(held, vals)
    d = D.Desc([A.m, A.n], [D.DENSE, D.SPARSE], D.CSIZE[dtype])
    d._m = ctypes.c_int32(A.m)
    d._level_arrays[0][0] = ctypes.cast(ctypes.pointer(d._m), ctypes.POINTER(ctypes.c_uint8))
    d.set_level(1, [p.addr, c.addr])
    d.set_vals(v.addr, A.nnz)
    return d


def dense_managed(held, host, dtype="f64"):
    arr = managed(held, host)
    d = D.Desc(list(host.shape), [D.DENSE] * host.ndim, D.CSIZE[dtype])
    d._sizes = [ctypes.c_int32(s) for s in host.shape]
    for k, s in enumerate(d._sizes):
        d._level_arrays[k][0] = ctypes.cast(ctypes.pointer(s), ctypes.POINTER(ctypes.c_uint8))
    d.set_vals(arr.addr, host.size)
    return d, arr


@pytest.fixture(scope="module")
def mat():
    A = W.csr_powerlaw(20000, 20000, seed=31, kmax=6000)
    assert np.diff(A.pos).max() > 1024  # long rows: MFMA splits them, atomics join the parts
    return A


def test_unified_alloc_is_managed(unified):
    a = managed(unified, np.arange(100, dtype=np.float64))
    attr = _lib.lib().taco_hip_should_use_unified_memory()
    assert attr == 1
    assert np.array_equal(a.read(), np.arange(100, dtype=np.float64))


@pytest.mark.parametrize("dtype,tol", [("f64", 1e-12), ("f32", 1e-5)])
@pytest.mark.parametrize("algo,items", [(_lib.NNZ_SPLIT, 8), (_lib.MERGE_PATH, 8),
                                        (_lib.THREAD_PER_ROW, 0)])
def test_unified_spmv_hand_written(unified, mat, dtype, tol, algo, items):
    A = mat
    x = W.uniform(A.n, 31, W.STREAM_X)
    npdt = np.float64 if dtype == "f64" else np.float32
    a = csr_managed(unified, A, dtype)
    xd, _ = dense_managed(unified, x.astype(npdt), dtype)
    yd, y = dense_managed(unified, np.full(A.m, np.nan, npdt), dtype)
    _lib.check(_lib.lib().taco_hip_spmv_compute(yd.ptr, a.ptr, xd.ptr,
                                                G.sched(algo, 256, items)))
    # the oracle on the values the kernel sees (rounded to f32 for f32)
    Ar = W.Csr(A.m, A.n, A.pos, A.crd, A.vals.astype(npdt).astype(np.float64))
    ye, cond = oracle.exact_spmv(Ar, x.astype(npdt).astype(np.float64))
    rep = oracle.error_report(y.read(), ye, cond, tol)
    assert rep["ok"], rep


def test_unified_spmv_slab_plan(unified, mat):
    A = mat
    x = W.uniform(A.n, 32, W.STREAM_X)
    a = csr_managed(unified, A)
    xd, _ = dense_managed(unified, x)
    yd, y = dense_managed(unified, np.full(A.m, np.nan))
    plan = _lib.SpmvPlan(a, 64, 4096)
    plan.compute(yd, xd, 0)
    ye, cond = oracle.exact_spmv(A, x)
    rep = oracle.error_report(y.read(), ye, cond, 1e-12)
    assert rep["ok"], rep
    del plan


@pytest.mark.parametrize("dtype,tol", [("f64", 1e-12), ("f32", 1e-5), ("bf16", 1e-5)])
def test_unified_spmm_atomic_schedules(unified, mat, dtype, tol):
    """nnz-split (f64, f32) and MFMA (bf16, whose long rows are split into
    parts joined by atomics) add into C with float atomics."""
    A = mat
    K = 32
    B = W.uniform(A.n * K, 33, W.STREAM_B).reshape(A.n, K)
    if dtype == "bf16":
        a = csr_managed(unified, A, "bf16")
        bd, _ = dense_managed(unified, G.bf16_bits(B).numpy(), "bf16")
        cd, C = dense_managed(unified, np.full((A.m, K), np.nan, np.float32), "f32")
        s = G.sched(_lib.MFMA)
        Ar = W.Csr(A.m, A.n, A.pos, A.crd, G.bf16_values(A.vals))
        Br = G.bf16_values(B)
    else:
        npdt = np.float64 if dtype == "f64" else np.float32
        a = csr_managed(unified, A, dtype)
        bd, _ = dense_managed(unified, B.astype(npdt), dtype)
        cd, C = dense_managed(unified, np.full((A.m, K), np.nan, npdt), dtype)
        s = G.sched(_lib.NNZ_SPLIT, 256, 8, 32)
        Ar = W.Csr(A.m, A.n, A.pos, A.crd, A.vals.astype(npdt).astype(np.float64))
        Br = B.astype(npdt).astype(np.float64)
    _lib.check(_lib.lib().taco_hip_spmm_compute(cd.ptr, a.ptr, bd.ptr, s))
    Ce, cond = oracle.exact_spmm(Ar, Br)
    rep = oracle.error_report(C.read().astype(np.float64), Ce, cond, tol)
    assert rep["ok"], rep


@pytest.mark.parametrize("name", codegen.UNIFIED_PREBUILT)
def test_unified_generated_modules(unified, mat, name):
    """Generated modules compiled under unified memory (compare-and-swap
    atomics) on managed operands and outputs."""
    A = mat
    op, fam, csize = codegen.STANDARD_SPECS[name][:3]
    k = codegen.Kernel(codegen.standard(name))
    assert "TACO_HIP_COHERENT_ATOMICS" in _lib.lib().taco_hip_get_default_compiler_flags().decode()
    a = csr_managed(unified, A)
    if op == _lib.OP_SPMV:
        x = W.uniform(A.n, 34, W.STREAM_X)
        xd, _ = dense_managed(unified, x)
        yd, y = dense_managed(unified, np.full(A.m, np.nan))
        k.compute(yd, a, xd)
        ye, cond = oracle.exact_spmv(A, x)
        got = y.read()
    else:
        K = 64
        B = W.uniform(A.n * K, 34, W.STREAM_B).reshape(A.n, K)
        bd, _ = dense_managed(unified, B)
        cd, C = dense_managed(unified, np.full((A.m, K), np.nan))
        k.compute(cd, a, bd)
        ye, cond = oracle.exact_spmm(A, B)
        got = C.read()
    rep = oracle.error_report(got, ye, cond, 1e-12)
    assert rep["ok"], (name, rep)
